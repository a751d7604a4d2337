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

struct PlannedPass {
     int k = 0;
     uint8_t pos[5] = {0, 0, 0, 0, 0};
     std::vector<cplx> m;  // 4^k entries, row-major
};

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

Batch::Batch(Prec prec, int device) : prec_(prec), device_(device)
{
     int ndev = 0;
     check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
     if (device < 0 || device >= ndev)
          throw std::runtime_error("Batch(): no HIP device " + std::to_string(device));
     DeviceScope ds(device_);
     check(hipStreamCreateWithFlags(&s1_, hipStreamNonBlocking), "hipStreamCreate");
     check(hipStreamCreateWithFlags(&s2_, hipStreamNonBlocking), "hipStreamCreate");
     check(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming), "hipEventCreate");
     check(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming), "hipEventCreate");
     check(hipEventCreate(&ev_t0_), "hipEventCreate");
     check(hipEventCreate(&ev_t1_), "hipEventCreate");
}

Batch::~Batch()
{
     (void)hipSetDevice(device_);
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

void Batch::set_plan(const std::vector<int>& k_by_n, int wg_max_qubits)
{
     for (int k: k_by_n)
          if (k < 0 || k > kMaxWidth) throw std::runtime_error("batch: fusion width must be 0..5");
     if (wg_max_qubits < 0 || wg_max_qubits > kMaxWgQubits)
          throw std::runtime_error("batch: wg_max_qubits must be 0.." + std::to_string(kMaxWgQubits));
     k_by_n_ = k_by_n;
     wg_max_qubits_ = wg_max_qubits;
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

void Batch::build()
{
     DeviceScope ds(device_);
     check(hipStreamSynchronize(s1_), "hipStreamSynchronize");
     release_device();
     stats_ = BatchStats();
     const size_t nc = circ_.size();
     if (nc == 0) throw std::runtime_error("batch: no circuits");
     const size_t ab = amp_bytes(prec_);

     std::vector<std::vector<PlannedPass>> plan(nc);
     parallel_for(nc, [&](size_t c) { plan[c] = plan_circuit(circ_[c].gates, fusion_width(circ_[c].n)); });

     // Classes and arena layout: layered states first (one zero fill), then
     // the workgroup-owned ones.
     std::vector<size_t> layer_c, wg_c;
     for (size_t c = 0; c < nc; ++c) (circ_[c].n > wg_max_qubits_ ? layer_c : wg_c).push_back(c);
     uint64_t off = 0;
     for (size_t c: layer_c) {
          circ_[c].state_off = off;
          off += uint64_t(1) << circ_[c].n;
     }
     const uint64_t layer_amps = off;
     for (size_t c: wg_c) {
          circ_[c].state_off = off;
          off += uint64_t(1) << circ_[c].n;
     }
     const uint64_t arena_amps = off;

     // Matrices, concatenated in the state's precision.
     std::vector<std::vector<uint64_t>> mat_off(nc);
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
     }
     std::vector<char> h_mats(std::max<uint64_t>(mat_entries, 1) * ab);
     parallel_for(nc, [&](size_t c) {
          for (size_t p = 0; p < plan[c].size(); ++p) {
               const auto& m = plan[c][p].m;
               if (prec_ == Prec::C128) {
                    double* d = reinterpret_cast<double*>(h_mats.data()) + 2 * mat_off[c][p];
                    for (size_t i = 0; i < m.size(); ++i) {
                         d[2 * i] = m[i].real();
                         d[2 * i + 1] = m[i].imag();
                    }
               }
               else {
                    float* d = reinterpret_cast<float*>(h_mats.data()) + 2 * mat_off[c][p];
                    for (size_t i = 0; i < m.size(); ++i) {
                         d[2 * i] = static_cast<float>(m[i].real());
                         d[2 * i + 1] = static_cast<float>(m[i].imag());
                    }
               }
          }
     });

     // Workgroup-owned circuits: descriptors and a longest-first assignment
     // to workgroups by estimated cost (LDS-resident passes are cheaper).
     int cus = 0;
     check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_),
           "hipDeviceGetAttribute");
     const uint64_t lds_amps = kBatchLdsBytes / ab;
     std::vector<BatchPass> h_pass;
     std::vector<BatchCirc> h_circ;
     std::vector<double> cost;
     int kmax_wg = 1;
     for (size_t c: wg_c) {
          BatchCirc bc{};
          bc.state_off = circ_[c].state_off;
          bc.n = static_cast<uint32_t>(circ_[c].n);
          bc.pass_begin = static_cast<uint32_t>(h_pass.size());
          const uint64_t N = uint64_t(1) << circ_[c].n;
          const bool in_lds = N <= lds_amps;
          double cst = static_cast<double>(N);
          for (size_t p = 0; p < plan[c].size(); ++p) {
               BatchPass bp{};
               bp.mat_off = mat_off[c][p];
               bp.k = static_cast<uint8_t>(plan[c][p].k);
               std::memcpy(bp.pos, plan[c][p].pos, 5);
               h_pass.push_back(bp);
               kmax_wg = std::max(kmax_wg, plan[c][p].k);
               cst += static_cast<double>(N) * ((1 << plan[c][p].k) + (in_lds ? 2 : 8));
          }
          bc.pass_end = static_cast<uint32_t>(h_pass.size());
          h_circ.push_back(bc);
          cost.push_back(cst);
          if (in_lds) ++stats_.lds_circuits;
     }
     stats_.wg_circuits = wg_c.size();
     const int n_wg = static_cast<int>(std::min<size_t>(static_cast<size_t>(std::max(cus, 1)), wg_c.size()));
     std::vector<uint32_t> h_wg_off(n_wg + 1, 0), h_wg_list;
     if (n_wg > 0) {
          std::vector<size_t> order(wg_c.size());
          for (size_t i = 0; i < order.size(); ++i) order[i] = i;
          std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
          using Slot = std::pair<double, int>;
          std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
          for (int w = 0; w < n_wg; ++w) heap.push({0.0, w});
          std::vector<std::vector<uint32_t>> lists(n_wg);
          for (size_t i: order) {
               Slot s = heap.top();
               heap.pop();
               lists[s.second].push_back(static_cast<uint32_t>(i));
               heap.push({s.first + cost[i], s.second});
          }
          for (int w = 0; w < n_wg; ++w) {
               h_wg_off[w + 1] = h_wg_off[w] + static_cast<uint32_t>(lists[w].size());
               h_wg_list.insert(h_wg_list.end(), lists[w].begin(), lists[w].end());
          }
     }
     stats_.wg_grid = n_wg;

     // Layered circuits: launch (L, k) holds the L-th pass of every layered
     // circuit whose L-th pass has width k.
     struct LayerLaunch {
          int k;
          size_t first, count;
          uint64_t blocks;
     };
     std::vector<BatchItem> h_items;
     std::vector<LayerLaunch> launches;
     size_t max_len = 0;
     for (size_t c: layer_c) max_len = std::max(max_len, plan[c].size());
     for (size_t L = 0; L < max_len; ++L) {
          for (int k = 1; k <= kMaxWidth; ++k) {
               LayerLaunch ll{k, h_items.size(), 0, 0};
               for (size_t c: layer_c) {
                    if (L >= plan[c].size() || plan[c][L].k != k) continue;
                    BatchItem it{};
                    it.state_off = circ_[c].state_off;
                    it.mat_off = mat_off[c][L];
                    it.block_start = ll.blocks;
                    it.log_groups = static_cast<uint32_t>(circ_[c].n - k);
                    std::memcpy(it.pos, plan[c][L].pos, 5);
                    const uint64_t groups = uint64_t(1) << it.log_groups;
                    ll.blocks += (groups + kBatchLayerBlock - 1) / kBatchLayerBlock;
                    h_items.push_back(it);
                    ++ll.count;
               }
               if (ll.count) launches.push_back(ll);
          }
     }
     stats_.layer_launches = launches.size();
     std::vector<uint64_t> h_ones;
     for (size_t c: layer_c) h_ones.push_back(circ_[c].state_off);

     // Device memory.
     stats_.circuits = nc;
     stats_.arena_bytes = arena_amps * ab;
     stats_.matrix_bytes = mat_entries * ab;
     size_t free_b = 0, total_b = 0;
     check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
     const size_t sz_pass = align_up(h_pass.size() * sizeof(BatchPass), 256);
     const size_t sz_circ = align_up(h_circ.size() * sizeof(BatchCirc), 256);
     const size_t sz_woff = align_up(h_wg_off.size() * sizeof(uint32_t), 256);
     const size_t sz_wl = align_up(h_wg_list.size() * sizeof(uint32_t), 256);
     const size_t sz_items = align_up(h_items.size() * sizeof(BatchItem), 256);
     const size_t sz_ones = align_up(h_ones.size() * sizeof(uint64_t), 256);
     const size_t desc_bytes = sz_pass + sz_circ + sz_woff + sz_wl + sz_items + sz_ones + 256;
     const size_t need = stats_.arena_bytes + h_mats.size() + desc_bytes;
     if (need + (size_t(1) << 30) > free_b) {
          std::ostringstream os;
          os << "batch: needs " << need << " bytes of device memory, " << free_b << " free";
          throw std::runtime_error(os.str());
     }
     check(hipMalloc(&d_arena_, std::max<size_t>(stats_.arena_bytes, 256)), "hipMalloc(arena)");
     check(hipMalloc(&d_mats_, h_mats.size()), "hipMalloc(matrices)");
     check(hipMalloc(&d_desc_, desc_bytes), "hipMalloc(descriptors)");
     check(hipMemcpy(d_mats_, h_mats.data(), h_mats.size(), hipMemcpyHostToDevice), "hipMemcpy(matrices)");
     char* base = static_cast<char*>(d_desc_);
     auto* d_pass = reinterpret_cast<BatchPass*>(base);
     auto* d_circ = reinterpret_cast<BatchCirc*>(base + sz_pass);
     auto* d_woff = reinterpret_cast<uint32_t*>(base + sz_pass + sz_circ);
     auto* d_wl = reinterpret_cast<uint32_t*>(base + sz_pass + sz_circ + sz_woff);
     auto* d_items = reinterpret_cast<BatchItem*>(base + sz_pass + sz_circ + sz_woff + sz_wl);
     auto* d_ones = reinterpret_cast<uint64_t*>(base + sz_pass + sz_circ + sz_woff + sz_wl + sz_items);
     auto up = [](void* d, const void* h, size_t bytes) {
          if (bytes) check(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice), "hipMemcpy(descriptors)");
     };
     up(d_pass, h_pass.data(), h_pass.size() * sizeof(BatchPass));
     up(d_circ, h_circ.data(), h_circ.size() * sizeof(BatchCirc));
     up(d_woff, h_wg_off.data(), h_wg_off.size() * sizeof(uint32_t));
     up(d_wl, h_wg_list.data(), h_wg_list.size() * sizeof(uint32_t));
     up(d_items, h_items.data(), h_items.size() * sizeof(BatchItem));
     up(d_ones, h_ones.data(), h_ones.size() * sizeof(uint64_t));
     plan.clear();

     // Capture: the persistent launch runs on a forked stream beside the
     // layered chain; both join before the graph ends.
     check(hipStreamBeginCapture(s1_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
     const bool fork = n_wg > 0 && !layer_c.empty();
     hipStream_t ws = fork ? s2_ : s1_;
     if (fork) {
          check(hipEventRecord(ev_fork_, s1_), "hipEventRecord");
          check(hipStreamWaitEvent(s2_, ev_fork_, 0), "hipStreamWaitEvent");
     }
     if (n_wg > 0)
          check(launch_batch_wg(d_arena_, prec_, d_mats_, d_pass, d_circ, d_woff, d_wl, n_wg,
                                kmax_wg, ws),
                "launch_batch_wg");
     if (fork) check(hipEventRecord(ev_join_, s2_), "hipEventRecord");
     if (!layer_c.empty()) {
          check(launch_fill(d_arena_, prec_, layer_amps, 0.0, 0.0, s1_), "launch_fill");
          check(launch_set_ones(d_arena_, prec_, d_ones, static_cast<int>(h_ones.size()), s1_),
                "launch_set_ones");
          for (const auto& ll: launches)
               check(launch_batch_layer(d_arena_, prec_, d_mats_, d_items + ll.first,
                                        static_cast<int>(ll.count), ll.k, ll.blocks, s1_),
                     "launch_batch_layer");
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
