// Host side of the batched many-circuit path.  See batch.hpp.
#include "batch.hpp"

#include <algorithm>
#include <cstring>
#include <functional>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "fusion.hpp"
#include "sched.hpp"

namespace hq {
namespace {

constexpr int kDefaultWidth = 4;
constexpr int kMaxWidth = 5;
constexpr int kMaxBatchQubits = 30;
constexpr int kMaxWgQubits = 24;  // workgroup-owned states use 32-bit indices

void check(hipError_t e, const char* what)
{
     if (e != hipSuccess) {
          std::ostringstream os;
          os << what << ": " << hipGetErrorString(e);
          throw std::runtime_error(os.str());
     }
}

struct DeviceScope {
     explicit DeviceScope(int dev) { check(hipSetDevice(dev), "hipSetDevice"); }
};

template <class F>
void parallel_for(size_t n, F&& f)
{
     size_t nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
     nt = std::min(nt, n);
     if (nt <= 1) {
          for (size_t i = 0; i < n; ++i) f(i);
          return;
     }
     std::vector<std::thread> th;
     std::vector<std::exception_ptr> err(nt);
     for (size_t t = 0; t < nt; ++t)
          th.emplace_back([&, t] {
               try {
                    for (size_t i = t; i < n; i += nt) f(i);
               } catch (...) {
                    err[t] = std::current_exception();
               }
          });
     for (auto& x: th) x.join();
     for (auto& e: err)
          if (e) std::rethrow_exception(e);
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

// Clusters of at most k qubits (lookahead packer), each fused into one dense
// matrix over its ascending qubits; controls enter the matrix as diag(I, U)
// on the control bits (fusion_mpi.hpp:167-226 folds them the same way).
std::vector<PlannedPass> plan_circuit(const std::vector<BatchGate>& gates, int k)
{
     std::vector<PlannedPass> out;
     if (gates.empty()) return out;
     GateList gl;
     for (const auto& g: gates) {
          gl.gate.push_back(g.targets);
          gl.ctrl.push_back(g.ctrls);
          gl.diag.push_back(false);
     }
     const std::vector<int> cluster_of = plan_clusters(gl, k, kPlanLookahead);
     const int nc = *std::max_element(cluster_of.begin(), cluster_of.end()) + 1;
     std::vector<std::vector<int>> members(nc);
     for (size_t i = 0; i < gates.size(); ++i) members[cluster_of[i]].push_back(static_cast<int>(i));
     out.reserve(nc);
     for (const auto& mem: members) {
          Fusion f;
          for (int gi: mem) {
               const BatchGate& g = gates[gi];
               const size_t dim = size_t(1) << g.targets.size();
               Mat m(dim);
               m.a = g.m;
               std::vector<int64_t> ids = g.targets;
               add_controls(m, ids, g.ctrls);
               f.insert(std::move(m), false, std::move(ids), {});
          }
          Mat M;
          std::vector<int64_t> ids, ctrls;
          bool diag = false;
          f.perform(M, ids, ctrls, diag);
          if (ids.size() > static_cast<size_t>(kMaxWidth) || !ctrls.empty())
               throw std::runtime_error("batch: a fused cluster spans more than 5 qubits");
          PlannedPass p;
          p.k = static_cast<int>(ids.size());
          for (int l = 0; l < p.k; ++l) p.pos[l] = static_cast<uint8_t>(ids[l]);
          p.m = std::move(M.a);
          out.push_back(std::move(p));
     }
     return out;
}

PlannedTile fuse_tile(const std::vector<BatchGate>& gates, const TilePlan& tp)
{
     PlannedTile t;
     int rank[64];
     for (int p = 0; p < 64; ++p) rank[p] = p < 6 ? p : -1;
     if (tp.hpos.size() != static_cast<size_t>(kTileHigh)) throw std::runtime_error("batch: bad tile plan");
     for (int j = 0; j < kTileHigh; ++j) {
          t.hpos[j] = static_cast<uint8_t>(tp.hpos[j]);
          rank[tp.hpos[j]] = 6 + j;
     }
     for (const auto& op: tp.ops) {
          Fusion f;
          for (int gi: op) {
               const BatchGate& g = gates[gi];
               Mat m(size_t(1) << g.targets.size());
               m.a = g.m;
               std::vector<int64_t> ids = g.targets;
               add_controls(m, ids, g.ctrls);
               f.insert(std::move(m), false, std::move(ids), {});
          }
          Mat M;
          std::vector<int64_t> ids, ctrls;
          bool diag = false;
          f.perform(M, ids, ctrls, diag);
          if (ids.size() > static_cast<size_t>(kMaxWidth) || !ctrls.empty())
               throw std::runtime_error("batch: a fused op spans more than 5 qubits");
          PlannedPass p;
          p.k = static_cast<int>(ids.size());
          for (int l = 0; l < p.k; ++l) {
               if (rank[ids[l]] < 0) throw std::runtime_error("batch: op outside its tile");
               p.pos[l] = static_cast<uint8_t>(rank[ids[l]]);
          }
          p.m = std::move(M.a);
          t.ops.push_back(std::move(p));
     }
     return t;
}

GateList gate_list(const std::vector<BatchGate>& gates)
{
     GateList gl;
     for (const auto& g: gates) {
          gl.gate.push_back(g.targets);
          gl.ctrl.push_back(g.ctrls);
          gl.diag.push_back(false);
     }
     return gl;
}

std::vector<PlannedTile> plan_circuit_tiles(const std::vector<BatchGate>& gates, int n, int op_width)
{
     std::vector<PlannedTile> out;
     if (gates.empty()) return out;
     for (const TilePlan& tp: plan_tile_passes(gate_list(gates), n, kTileHigh, op_width))
          out.push_back(fuse_tile(gates, tp));
     return out;
}


// Host-side layout of a wave-tile item (batch.hpp).
bool make_tile_item(int n, int k, const uint8_t* pos, Prec prec, BatchTileItem& it)
{
     const int wpb = batch_tile_wpb(k, prec);
     if (k < 1 || k > 5 || n - k < 6 + (wpb == 4 ? 2 : 1)) return false;
     uint64_t tmask = 0;
     for (int l = 0; l < k; ++l) tmask |= 1ull << pos[l];
     // lanes: the six lowest non-target positions
     uint64_t lanes = 0;
     int nb = 0;
     for (int p = 0; p < n && nb < 6; ++p)
          if (!((tmask >> p) & 1)) {
               lanes |= 1ull << p;
               ++nb;
          }
     const uint64_t vbits = lanes | tmask;
     if ((vbits & 63) != 63) return false;
     int vpos[64], rank_of[64], nv = 0;
     for (int p = 0; p < 64; ++p)
          if ((vbits >> p) & 1) {
               rank_of[p] = nv;
               vpos[nv++] = p;
          }
     if (nv != 6 + k) return false;
     nb = 0;
     for (int p = 0; p < 64; ++p)
          if ((lanes >> p) & 1) it.rank_n[nb++] = static_cast<uint8_t>(rank_of[p]);
     for (int l = 0; l < k; ++l) {
          it.rank_t[l] = static_cast<uint8_t>(rank_of[pos[l]]);
          it.ins[l] = pos[l];
     }
     it.nins = static_cast<uint8_t>(k);
     it.log_blocks = static_cast<uint8_t>(n - k - 6 - (wpb == 4 ? 2 : 1));
     for (int i = 0; i < 32; ++i) {
          uint64_t o = 0;
          if (i < (1 << k))
               for (int q = 0; q < k; ++q)
                    if ((i >> q) & 1) o |= 1ull << vpos[6 + q];
          it.hoff[i] = static_cast<uint32_t>(o);
     }
     return true;
}

Batch::Batch(Prec prec, int device) : prec_(prec), device_(device)
{
     if (device < 0) throw std::runtime_error("Batch(): device must be >= 0");
}

void Batch::ensure_device()
{
     if (device_ready_) return;
     int ndev = 0;
     check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
     if (device_ >= ndev) throw std::runtime_error("Batch(): no HIP device " + std::to_string(device_));
     DeviceScope ds(device_);
     check(hipStreamCreateWithFlags(&s1_, hipStreamNonBlocking), "hipStreamCreate");
     check(hipStreamCreateWithFlags(&s2_, hipStreamNonBlocking), "hipStreamCreate");
     check(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming), "hipEventCreate");
     check(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming), "hipEventCreate");
     check(hipEventCreate(&ev_t0_), "hipEventCreate");
     check(hipEventCreate(&ev_t1_), "hipEventCreate");
     device_ready_ = true;
}

Batch::~Batch()
{
     if (!device_ready_) return;
     (void)hipSetDevice(device_);
     (void)hipStreamSynchronize(s1_);
     release_device();
     (void)hipEventDestroy(ev_fork_);
     (void)hipEventDestroy(ev_join_);
     (void)hipEventDestroy(ev_t0_);
     (void)hipEventDestroy(ev_t1_);
     (void)hipStreamDestroy(s1_);
     (void)hipStreamDestroy(s2_);
}

void Batch::release_device()
{
     if (exec_) (void)hipGraphExecDestroy(exec_);
     if (graph_) (void)hipGraphDestroy(graph_);
     exec_ = nullptr;
     graph_ = nullptr;
     for (void** p: {&d_arena_, &d_mats_, &d_desc_}) {
          if (*p) (void)hipFree(*p);
          *p = nullptr;
     }
     built_ = false;
}

int64_t Batch::add_circuit(int n, std::vector<BatchGate> gates)
{
     if (n < 1 || n > kMaxBatchQubits)
          throw std::runtime_error("batch: circuits must have 1.." + std::to_string(kMaxBatchQubits) +
                                   " qubits, got " + std::to_string(n));
     for (size_t gi = 0; gi < gates.size(); ++gi) {
          const auto& g = gates[gi];
          const std::string where = "batch: gate " + std::to_string(gi) + ": ";
          if (g.targets.empty()) throw std::runtime_error(where + "no target qubit");
          if (g.targets.size() + g.ctrls.size() > static_cast<size_t>(kMaxWidth))
               throw std::runtime_error(where + "more than 5 qubits with its controls");
          const size_t dim = size_t(1) << g.targets.size();
          if (g.m.size() != dim * dim) throw std::runtime_error(where + "matrix size does not match targets");
          std::vector<int64_t> all = g.targets;
          all.insert(all.end(), g.ctrls.begin(), g.ctrls.end());
          for (auto q: all)
               if (q < 0 || q >= n) throw std::runtime_error(where + "qubit id out of range");
          std::sort(all.begin(), all.end());
          if (std::adjacent_find(all.begin(), all.end()) != all.end())
               throw std::runtime_error(where + "repeated qubit");
     }
     circ_.push_back(Circ{n, std::move(gates), 0});
     built_ = false;
     return static_cast<int64_t>(circ_.size()) - 1;
}

int64_t Batch::add_copies(int64_t c, int64_t count)
{
     if (c < 0 || c >= num_circuits()) throw std::runtime_error("batch: no circuit " + std::to_string(c));
     if (count < 0) throw std::runtime_error("batch: negative copy count");
     const int64_t first = num_circuits();
     circ_.reserve(circ_.size() + static_cast<size_t>(count));
     for (int64_t i = 0; i < count; ++i) circ_.push_back(Circ{circ_[c].n, circ_[c].gates, 0});
     built_ = false;
     return first;
}

void Batch::set_plan(const std::vector<int>& k_by_n, int wg_max_qubits, int tile_min_qubits,
                     int prog_min_qubits)
{
     if (tile_min_qubits < 0 || tile_min_qubits > 31)
          throw std::runtime_error("batch: tile_min_qubits must be 0..31");
     if (prog_min_qubits < 0 || prog_min_qubits > 31)
          throw std::runtime_error("batch: prog_min_qubits must be 0..31");
     for (int k: k_by_n)
          if (k < 0 || k > kMaxWidth) throw std::runtime_error("batch: fusion width must be 0..5");
     if (wg_max_qubits < 0 || wg_max_qubits > kMaxWgQubits)
          throw std::runtime_error("batch: wg_max_qubits must be 0.." + std::to_string(kMaxWgQubits));
     k_by_n_ = k_by_n;
     wg_max_qubits_ = wg_max_qubits;
     tile_min_qubits_ = tile_min_qubits;
     prog_min_qubits_ = prog_min_qubits;
     built_ = false;
}

int Batch::fusion_width(int n) const
{
     int k = (n >= 0 && static_cast<size_t>(n) < k_by_n_.size()) ? k_by_n_[n] : 0;
     return k > 0 ? k : kDefaultWidth;
}

int Batch::num_qubits(int64_t c) const
{
     if (c < 0 || c >= num_circuits()) throw std::runtime_error("batch: no circuit " + std::to_string(c));
     return circ_[c].n;
}

// Host plan: every circuit cut into clusters and fused, the arena laid out
// (layered states first, one zero fill), the workgroup lists and the layered
// launches.
Batch::HostPlan Batch::plan_host(int workgroups)
{
     HostPlan hp;
     stats_ = BatchStats();
     const size_t nc = circ_.size();
     if (nc == 0) throw std::runtime_error("batch: no circuits");
     const size_t ab = amp_bytes(prec_);

     std::vector<size_t> layer_c, wg_c;
     for (size_t c = 0; c < nc; ++c) (circ_[c].n > wg_max_qubits_ ? layer_c : wg_c).push_back(c);
     const int prog_min = std::max(prog_min_qubits_, kTileBits);
     std::vector<uint8_t> is_prog(nc, 0);
     for (size_t c: layer_c) is_prog[c] = circ_[c].n >= prog_min;
     std::vector<std::vector<PlannedPass>> plan(nc);
     std::vector<std::vector<PlannedTile>> tiles(nc);
     parallel_for(nc, [&](size_t c) {
          if (is_prog[c]) tiles[c] = plan_circuit_tiles(circ_[c].gates, circ_[c].n, fusion_width(circ_[c].n));
          else plan[c] = plan_circuit(circ_[c].gates, fusion_width(circ_[c].n));
     });
     uint64_t off = 0;
     for (size_t c: layer_c) {
          circ_[c].state_off = off;
          off += uint64_t(1) << circ_[c].n;
     }
     hp.layer_amps = off;
     for (size_t c: wg_c) {
          circ_[c].state_off = off;
          off += uint64_t(1) << circ_[c].n;
     }
     hp.arena_amps = off;

     // Matrices, concatenated in the state's precision: each circuit's fused
     // passes, then its tile ops.
     std::vector<std::vector<uint64_t>> mat_off(nc), tmat_off(nc);
     uint64_t mat_entries = 0;
     for (size_t c = 0; c < nc; ++c) {
          mat_off[c].resize(plan[c].size());
          for (size_t p = 0; p < plan[c].size(); ++p) {
               mat_off[c][p] = mat_entries;
               mat_entries += plan[c][p].m.size();
               stats_.amp_passes += uint64_t(1) << circ_[c].n;
               stats_.kmax = std::max(stats_.kmax, plan[c][p].k);
          }
          stats_.passes += plan[c].size();
          for (const auto& t: tiles[c]) {
               for (const auto& op: t.ops) {
                    tmat_off[c].push_back(mat_entries);
                    mat_entries += op.m.size();
                    stats_.kmax = std::max(stats_.kmax, op.k);
                    if (op.k >= 3) stats_.tile_mfma_flops += uint64_t(8) << (op.k + circ_[c].n);
               }
               stats_.amp_passes += uint64_t(1) << circ_[c].n;
               stats_.tile_ops += t.ops.size();
          }
          stats_.passes += tiles[c].size();
          stats_.tile_passes += tiles[c].size();
     }
     hp.mat_entries = mat_entries;
     hp.mats.resize(std::max<uint64_t>(mat_entries, 1) * ab);
     std::vector<char>& h_mats = hp.mats;
     parallel_for(nc, [&](size_t c) {
          size_t q = 0;
          for (const auto& t: tiles[c])
               for (const auto& op: t.ops)
                    put_tile_matrix(h_mats.data() + tmat_off[c][q++] * ab, op.m.data(), op.k, prec_);
          for (size_t p = 0; p < plan[c].size(); ++p) {
               const auto& m = plan[c][p].m;
               const uint64_t moff = mat_off[c][p];
               if (prec_ == Prec::C128) {
                    double* d = reinterpret_cast<double*>(h_mats.data()) + 2 * moff;
                    for (size_t i = 0; i < m.size(); ++i) {
                         d[2 * i] = m[i].real();
                         d[2 * i + 1] = m[i].imag();
                    }
               }
               else {
                    float* d = reinterpret_cast<float*>(h_mats.data()) + 2 * moff;
                    for (size_t i = 0; i < m.size(); ++i) {
                         d[2 * i] = static_cast<float>(m[i].real());
                         d[2 * i + 1] = static_cast<float>(m[i].imag());
                    }
               }
          }
     });

     // Workgroup-owned circuits: descriptors and a longest-first assignment
     // to workgroups by estimated cost (LDS-resident passes are cheaper).
     const uint64_t lds_amps = kBatchLdsBytes / ab;
     std::vector<double> cost;
     for (size_t c: wg_c) {
          BatchCirc bc{};
          bc.state_off = circ_[c].state_off;
          bc.n = static_cast<uint32_t>(circ_[c].n);
          bc.pass_begin = static_cast<uint32_t>(hp.pass.size());
          const uint64_t N = uint64_t(1) << circ_[c].n;
          const bool in_lds = N <= lds_amps;
          double cst = static_cast<double>(N);
          for (size_t p = 0; p < plan[c].size(); ++p) {
               BatchPass bp{};
               bp.mat_off = mat_off[c][p];
               bp.k = static_cast<uint8_t>(plan[c][p].k);
               std::memcpy(bp.pos, plan[c][p].pos, 5);
               hp.pass.push_back(bp);
               hp.kmax_wg = std::max(hp.kmax_wg, plan[c][p].k);
               cst += static_cast<double>(N) * ((1 << plan[c][p].k) + (in_lds ? 2 : 8));
          }
          bc.pass_end = static_cast<uint32_t>(hp.pass.size());
          hp.circ.push_back(bc);
          cost.push_back(cst);
          if (in_lds) ++stats_.lds_circuits;
     }
     stats_.wg_circuits = wg_c.size();
     hp.n_wg = static_cast<int>(std::min<size_t>(static_cast<size_t>(std::max(workgroups, 1)), wg_c.size()));
     const int n_wg = hp.n_wg;
     hp.wg_off.assign(n_wg + 1, 0);
     if (n_wg > 0) {
          std::vector<size_t> order(wg_c.size());
          for (size_t i = 0; i < order.size(); ++i) order[i] = i;
          std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
          using Slot = std::pair<double, int>;
          std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
          for (int w = 0; w < n_wg; ++w) heap.push({0.0, w});
          std::vector<std::vector<uint32_t>> lists(n_wg);
          for (size_t i: order) {
               Slot sl = heap.top();
               heap.pop();
               lists[sl.second].push_back(static_cast<uint32_t>(i));
               heap.push({sl.first + cost[i], sl.second});
          }
          for (int w = 0; w < n_wg; ++w) {
               hp.wg_off[w + 1] = hp.wg_off[w] + static_cast<uint32_t>(lists[w].size());
               hp.wg_list.insert(hp.wg_list.end(), lists[w].begin(), lists[w].end());
          }
     }
     stats_.wg_grid = n_wg;
     // Layered circuits: launch (L, k, form) holds the L-th pass of every
     // layered circuit whose L-th pass has width k and takes that form.  The
     // wave-tile form is used unless every target is at bit >= 20 (chunks
     // 2^20 amplitudes apart) or, for k = 5, bit 0 (bits 0..1 in complex64)
     // is not a target (a 32 KiB-per-wave tile halves the occupancy).
     std::vector<BatchItem>& h_items = hp.items;
     std::vector<BatchTileItem>& h_titems = hp.titems;
     std::vector<LayerLaunch>& launches = hp.launches;
     size_t max_len = 0;
     for (size_t c: layer_c) max_len = std::max(max_len, plan[c].size());
     for (size_t L = 0; L < max_len; ++L) {
          for (int k = 1; k <= kMaxWidth; ++k) {
               LayerLaunch ld{k, kDirect, h_items.size(), 0, 0};
               LayerLaunch lt{k, kTileForm, h_titems.size(), 0, 0};
               for (size_t c: layer_c) {
                    if (L >= plan[c].size() || plan[c][L].k != k) continue;
                    const PlannedPass& pp = plan[c][L];
                    const int n = circ_[c].n;
                    BatchTileItem ti{};
                    const bool tile_pays = n >= tile_min_qubits_ && pp.pos[0] < get_tile_high_target_for(k) &&
                                           (k < 5 || pp.pos[0] <= (prec_ == Prec::C128 ? 0 : 1));
                    if (tile_pays && make_tile_item(n, k, pp.pos, prec_, ti)) {
                         ti.state_off = circ_[c].state_off;
                         ti.mat_off = mat_off[c][L];
                         ti.block_start = lt.blocks;
                         lt.blocks += (uint64_t(1) << (n - k)) / (64ull * batch_tile_wpb(k, prec_));
                         h_titems.push_back(ti);
                         hp.titem_n.push_back(n);
                         ++lt.count;
                         continue;
                    }
                    BatchItem it{};
                    it.state_off = circ_[c].state_off;
                    it.mat_off = mat_off[c][L];
                    it.block_start = ld.blocks;
                    it.log_groups = static_cast<uint32_t>(n - k);
                    std::memcpy(it.pos, pp.pos, 5);
                    const uint64_t groups = uint64_t(1) << it.log_groups;
                    ld.blocks += (groups + kBatchLayerBlock - 1) / kBatchLayerBlock;
                    h_items.push_back(it);
                    hp.item_n.push_back(n);
                    ++ld.count;
               }
               if (ld.count) launches.push_back(ld);
               if (lt.count) launches.push_back(lt);
          }
     }
     // Tile-program circuits: launch L holds the L-th tile pass of each.
     size_t max_tiles = 0;
     for (size_t c: layer_c) max_tiles = std::max(max_tiles, tiles[c].size());
     for (size_t L = 0; L < max_tiles; ++L) {
          LayerLaunch lp{1, kTileProgram, hp.progs.size(), 0, 0};
          for (size_t c: layer_c) {
               if (!is_prog[c] || L >= tiles[c].size()) continue;
               const PlannedTile& t = tiles[c][L];
               size_t q0 = 0;
               for (size_t l = 0; l < L; ++l) q0 += tiles[c][l].ops.size();
               BatchTileProg it{};
               it.state_off = circ_[c].state_off;
               it.block_start = lp.blocks;
               it.op_first = static_cast<uint32_t>(hp.tops.size());
               it.nops = static_cast<uint16_t>(t.ops.size());
               it.log_blocks = static_cast<uint8_t>(circ_[c].n - kTileBits);
               std::memcpy(it.hpos, t.hpos, kTileHigh);
               for (size_t o = 0; o < t.ops.size(); ++o) {
                    TileOp to{};
                    to.mat_off = tmat_off[c][q0 + o];
                    to.k = static_cast<uint8_t>(t.ops[o].k);
                    std::memcpy(to.rpos, t.ops[o].pos, 5);
                    lp.k = std::max(lp.k, t.ops[o].k);
                    hp.tops.push_back(to);
               }
               if (t.ops.size() > 65535) throw std::runtime_error("batch: too many ops in a tile pass");
               lp.blocks += uint64_t(1) << it.log_blocks;
               hp.progs.push_back(it);
               hp.prog_n.push_back(circ_[c].n);
               ++lp.count;
          }
          if (lp.count) launches.push_back(lp);
     }
     stats_.layer_launches = launches.size();
     for (const auto& ll: launches)
          if (ll.blocks > 65536 && ll.blocks % 65536 != 0) ++stats_.padded_launches;
     for (size_t c: layer_c) hp.ones.push_back(circ_[c].state_off);
     stats_.circuits = nc;
     stats_.arena_bytes = hp.arena_amps * ab;
     stats_.matrix_bytes = mat_entries * ab;
     return hp;
}


namespace {
uint64_t insert_zero_bits(uint64_t t, const uint8_t* pos, int k)
{
     for (int l = 0; l < k; ++l) {
          const uint64_t lo = t & ((uint64_t(1) << pos[l]) - 1);
          t = ((t ^ lo) << 1) | lo;
     }
     return t;
}
[[noreturn]] void plan_error(const std::string& what)
{
     throw std::runtime_error("batch: internal plan check failed: " + what);
}
}  // namespace

// Bounds of every launch against the buffers it touches: the kernels map
// blocks to items through block_start, so the items of a launch must tile
// [0, blocks) exactly and every address an item's blocks form must lie in
// its state; matrix reads must lie in the matrix array.
void Batch::check_plan(const HostPlan& hp) const
{
     auto state_ok = [&](uint64_t off, int n) { return n >= 1 && n <= 30 && off + (uint64_t(1) << n) <= hp.arena_amps; };
     auto mat_ok = [&](uint64_t off, int k) { return k >= 1 && k <= 5 && off + (uint64_t(1) << (2 * k)) <= hp.mat_entries; };
     auto pos_ok = [&](const uint8_t* pos, int k, int n) {
          for (int l = 0; l < k; ++l)
               if (pos[l] >= n || (l && pos[l] <= pos[l - 1])) return false;
          return true;
     };
     // workgroup-owned circuits
     if (hp.wg_off.size() != static_cast<size_t>(hp.n_wg) + 1 || hp.wg_off.back() != hp.wg_list.size() ||
         hp.wg_list.size() != hp.circ.size())
          plan_error("workgroup lists");
     for (int w = 0; w < hp.n_wg; ++w)
          if (hp.wg_off[w + 1] < hp.wg_off[w]) plan_error("workgroup offsets");
     for (uint32_t ci: hp.wg_list)
          if (ci >= hp.circ.size()) plan_error("workgroup list entry");
     for (const BatchCirc& c: hp.circ) {
          if (static_cast<int>(c.n) > kMaxWgQubits || !state_ok(c.state_off, static_cast<int>(c.n)))
               plan_error("workgroup state");
          if (c.pass_end < c.pass_begin || c.pass_end > hp.pass.size()) plan_error("workgroup passes");
          for (uint32_t p = c.pass_begin; p < c.pass_end; ++p) {
               const BatchPass& bp = hp.pass[p];
               if (bp.k > hp.kmax_wg || bp.k > static_cast<int>(c.n) || !mat_ok(bp.mat_off, bp.k) ||
                   !pos_ok(bp.pos, bp.k, static_cast<int>(c.n)))
                    plan_error("workgroup pass");
               }
     }
     if (hp.kmax_wg < 1 || hp.kmax_wg > 5) plan_error("workgroup width");
     // layered launches
     for (const LayerLaunch& ll: hp.launches) {
          uint64_t next = 0;
          for (size_t i = ll.first; i < ll.first + ll.count; ++i) {
               if (ll.kind == kTileProgram) {
                    if (i >= hp.progs.size()) plan_error("tile program index");
                    const BatchTileProg& it = hp.progs[i];
                    const int n = hp.prog_n[i];
                    if (n < kTileBits || it.log_blocks != n - kTileBits || it.block_start != next)
                         plan_error("tile program blocks");
                    next += uint64_t(1) << it.log_blocks;
                    if (!state_ok(it.state_off, n)) plan_error("tile program state");
                    for (int j = 0; j < kTileHigh; ++j)
                         if (it.hpos[j] < 6 || it.hpos[j] >= n || (j && it.hpos[j] <= it.hpos[j - 1]))
                              plan_error("tile program bits");
                    if (static_cast<size_t>(it.op_first) + it.nops > hp.tops.size()) plan_error("tile program ops");
                    for (size_t o = it.op_first; o < it.op_first + it.nops; ++o) {
                         const TileOp& to = hp.tops[o];
                         if (to.k > ll.k || !mat_ok(to.mat_off, to.k) || !pos_ok(to.rpos, to.k, kTileBits))
                              plan_error("tile program op");
                    }
                    continue;
               }
               if (ll.kind == kTileForm) {
                    if (i >= hp.titems.size()) plan_error("tile item index");
                    const BatchTileItem& it = hp.titems[i];
                    const int n = hp.titem_n[i];
                    const int wpb = batch_tile_wpb(ll.k, prec_);
                    const uint64_t nb = (uint64_t(1) << (n - ll.k)) / (64ull * wpb);
                    if (it.block_start != next || nb == 0 || nb * 64ull * wpb != (uint64_t(1) << (n - ll.k)) ||
                        (uint64_t(1) << it.log_blocks) != nb)
                         plan_error("tile item blocks");
                    next += nb;
                    if (it.nins != ll.k || !pos_ok(it.ins, ll.k, n) || !state_ok(it.state_off, n) ||
                        !mat_ok(it.mat_off, ll.k))
                         plan_error("tile item");
                    const uint64_t last = insert_zero_bits((nb * wpb - 1) << 6, it.ins, ll.k) + 63 +
                                          it.hoff[(1 << ll.k) - 1];
                    if (last >= (uint64_t(1) << n)) plan_error("tile item address");
               }
               else {
                    if (i >= hp.items.size()) plan_error("item index");
                    const BatchItem& it = hp.items[i];
                    const int n = hp.item_n[i];
                    if (it.log_groups != static_cast<uint32_t>(n - ll.k)) plan_error("item groups");
                    const uint64_t groups = uint64_t(1) << it.log_groups;
                    const uint64_t nb = (groups + kBatchLayerBlock - 1) / kBatchLayerBlock;
                    if (it.block_start != next) plan_error("item blocks");
                    next += nb;
                    if (!pos_ok(it.pos, ll.k, n) || !state_ok(it.state_off, n) || !mat_ok(it.mat_off, ll.k))
                         plan_error("item");
                    uint64_t span = 0;
                    for (int l = 0; l < ll.k; ++l) span |= uint64_t(1) << it.pos[l];
                    if (insert_zero_bits(groups - 1, it.pos, ll.k) + span >= (uint64_t(1) << n))
                         plan_error("item address");
               }
          }
          if (next != ll.blocks) plan_error("launch blocks");
     }
     for (uint64_t o: hp.ones)
          if (o >= hp.layer_amps) plan_error("initial amplitude offset");
}

void Batch::prepare()
{
     HostPlan hp = plan_host(kDefaultWorkgroups);
     check_plan(hp);
}

void Batch::build()
{
     ensure_device();
     DeviceScope ds(device_);
     check(hipStreamSynchronize(s1_), "hipStreamSynchronize");
     release_device();
     int cus = 0;
     check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_),
           "hipDeviceGetAttribute");
     HostPlan hp = plan_host(cus);
     check_plan(hp);
     const int n_wg = hp.n_wg;
     const bool has_layer = hp.layer_amps > 0;
     // Device memory.
     size_t free_b = 0, total_b = 0;
     check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
     const size_t sz_pass = align_up(hp.pass.size() * sizeof(BatchPass), 256);
     const size_t sz_circ = align_up(hp.circ.size() * sizeof(BatchCirc), 256);
     const size_t sz_woff = align_up(hp.wg_off.size() * sizeof(uint32_t), 256);
     const size_t sz_wl = align_up(hp.wg_list.size() * sizeof(uint32_t), 256);
     const size_t sz_items = align_up(hp.items.size() * sizeof(BatchItem), 256);
     const size_t sz_titems = align_up(hp.titems.size() * sizeof(BatchTileItem), 256);
     const size_t sz_ones = align_up(hp.ones.size() * sizeof(uint64_t), 256);
     const size_t sz_progs = align_up(hp.progs.size() * sizeof(BatchTileProg), 256);
     const size_t sz_tops = align_up(hp.tops.size() * sizeof(TileOp), 256);
     const size_t desc_bytes = sz_pass + sz_circ + sz_woff + sz_wl + sz_items + sz_titems + sz_ones + sz_progs +
                               sz_tops + 256;
     const size_t need = stats_.arena_bytes + hp.mats.size() + desc_bytes;
     if (need + (size_t(1) << 30) > free_b) {
          std::ostringstream os;
          os << "batch: needs " << need << " bytes of device memory, " << free_b << " free";
          throw std::runtime_error(os.str());
     }
     check(hipMalloc(&d_arena_, std::max<size_t>(stats_.arena_bytes, 256)), "hipMalloc(arena)");
     check(hipMalloc(&d_mats_, hp.mats.size()), "hipMalloc(matrices)");
     check(hipMalloc(&d_desc_, desc_bytes), "hipMalloc(descriptors)");
     check(hipMemcpy(d_mats_, hp.mats.data(), hp.mats.size(), hipMemcpyHostToDevice), "hipMemcpy(matrices)");
     char* base = static_cast<char*>(d_desc_);
     auto* d_pass = reinterpret_cast<BatchPass*>(base);
     auto* d_circ = reinterpret_cast<BatchCirc*>(base + sz_pass);
     auto* d_woff = reinterpret_cast<uint32_t*>(base + sz_pass + sz_circ);
     auto* d_wl = reinterpret_cast<uint32_t*>(base + sz_pass + sz_circ + sz_woff);
     auto* d_items = reinterpret_cast<BatchItem*>(base + sz_pass + sz_circ + sz_woff + sz_wl);
     auto* d_titems = reinterpret_cast<BatchTileItem*>(base + sz_pass + sz_circ + sz_woff + sz_wl + sz_items);
     auto* d_ones = reinterpret_cast<uint64_t*>(base + sz_pass + sz_circ + sz_woff + sz_wl + sz_items + sz_titems);
     auto* d_progs = reinterpret_cast<BatchTileProg*>(reinterpret_cast<char*>(d_ones) + sz_ones);
     auto* d_tops = reinterpret_cast<TileOp*>(reinterpret_cast<char*>(d_progs) + sz_progs);
     auto up = [](void* d, const void* h, size_t bytes) {
          if (bytes) check(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice), "hipMemcpy(descriptors)");
     };
     up(d_pass, hp.pass.data(), hp.pass.size() * sizeof(BatchPass));
     up(d_circ, hp.circ.data(), hp.circ.size() * sizeof(BatchCirc));
     up(d_woff, hp.wg_off.data(), hp.wg_off.size() * sizeof(uint32_t));
     up(d_wl, hp.wg_list.data(), hp.wg_list.size() * sizeof(uint32_t));
     up(d_items, hp.items.data(), hp.items.size() * sizeof(BatchItem));
     up(d_titems, hp.titems.data(), hp.titems.size() * sizeof(BatchTileItem));
     up(d_ones, hp.ones.data(), hp.ones.size() * sizeof(uint64_t));
     up(d_progs, hp.progs.data(), hp.progs.size() * sizeof(BatchTileProg));
     up(d_tops, hp.tops.data(), hp.tops.size() * sizeof(TileOp));

     // Capture: the persistent launch runs on a forked stream beside the
     // layered chain; both join before the graph ends.
     check(hipStreamBeginCapture(s1_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
     const bool fork = n_wg > 0 && has_layer;
     hipStream_t ws = fork ? s2_ : s1_;
     if (fork) {
          check(hipEventRecord(ev_fork_, s1_), "hipEventRecord");
          check(hipStreamWaitEvent(s2_, ev_fork_, 0), "hipStreamWaitEvent");
     }
     if (n_wg > 0)
          check(launch_batch_wg(d_arena_, prec_, d_mats_, d_pass, d_circ, d_woff, d_wl, n_wg,
                                hp.kmax_wg, ws),
                "launch_batch_wg");
     if (fork) check(hipEventRecord(ev_join_, s2_), "hipEventRecord");
     if (has_layer) {
          check(launch_fill(d_arena_, prec_, hp.layer_amps, 0.0, 0.0, s1_), "launch_fill");
          check(launch_set_ones(d_arena_, prec_, d_ones, static_cast<int>(hp.ones.size()), s1_),
                "launch_set_ones");
          for (const auto& ll: hp.launches) {
               if (ll.kind == kTileProgram)
                    check(launch_batch_tprog(d_arena_, prec_, d_mats_, d_progs + ll.first,
                                             static_cast<int>(ll.count), d_tops, ll.k, ll.blocks, s1_),
                          "launch_batch_tprog");
               else if (ll.kind == kTileForm)
                    check(launch_batch_tile(d_arena_, prec_, d_mats_, d_titems + ll.first,
                                            static_cast<int>(ll.count), ll.k, ll.blocks, s1_),
                          "launch_batch_tile");
               else
                    check(launch_batch_layer(d_arena_, prec_, d_mats_, d_items + ll.first,
                                             static_cast<int>(ll.count), ll.k, ll.blocks, s1_),
                          "launch_batch_layer");
          }
     }
     if (fork) check(hipStreamWaitEvent(s1_, ev_join_, 0), "hipStreamWaitEvent");
     check(hipStreamEndCapture(s1_, &graph_), "hipStreamEndCapture");
     check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
     built_ = true;
}

double Batch::run()
{
     if (!built_) throw std::runtime_error("batch: build() first");
     DeviceScope ds(device_);
     check(hipEventRecord(ev_t0_, s1_), "hipEventRecord");
     check(hipGraphLaunch(exec_, s1_), "hipGraphLaunch");
     check(hipEventRecord(ev_t1_, s1_), "hipEventRecord");
     check(hipEventSynchronize(ev_t1_), "hipEventSynchronize");
     float ms = 0;
     check(hipEventElapsedTime(&ms, ev_t0_, ev_t1_), "hipEventElapsedTime");
     return ms;
}

void Batch::get_state(int64_t c, std::complex<double>* out) const
{
     if (!built_) throw std::runtime_error("batch: build() first");
     const int n = num_qubits(c);
     DeviceScope ds(device_);
     const uint64_t N = uint64_t(1) << n;
     const size_t ab = amp_bytes(prec_);
     const char* src = static_cast<const char*>(d_arena_) + circ_[c].state_off * ab;
     if (prec_ == Prec::C128) {
          check(hipMemcpy(out, src, N * ab, hipMemcpyDeviceToHost), "hipMemcpy(state)");
          return;
     }
     std::vector<float> tmp(2 * N);
     check(hipMemcpy(tmp.data(), src, N * ab, hipMemcpyDeviceToHost), "hipMemcpy(state)");
     for (uint64_t i = 0; i < N; ++i) out[i] = std::complex<double>(tmp[2 * i], tmp[2 * i + 1]);
}

}  // namespace hq
