// Host side of the request loop.  See serve.hpp.
#include "serve.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "fusion.hpp"
#include "sched.hpp"

namespace hq {
namespace {

constexpr int kDefaultWidth = 4;
constexpr int kMaxWidth = 5;
constexpr int kMaxServeQubits = 30;

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

double now_s()
{
     return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Blocks of one item array laid out back to back (block_start ascending).
template <typename Item>
struct ItemListT {
     std::vector<Item> items;
     uint64_t blocks = 0;
     void add(Item it, uint64_t nblocks)
     {
          it.block_start = blocks;
          blocks += nblocks;
          items.push_back(it);
     }
};
using ItemList = ItemListT<BatchItem>;
using TileList = ItemListT<BatchTileItem>;

// Pass l of a step: the wave-tile items by width and the direct-form rest
// (widths mixed).
struct StepLayer {
     ItemList direct;
     TileList tile[6];
     ItemListT<BatchTileProg> progs;
     int prog_kmax = 1;
};

}  // namespace

// ---------------------------------------------------------------------------
Server::Server(Prec prec, int device, uint64_t state_bytes, uint64_t matrix_bytes,
               int passes_per_step, int max_active)
    : prec_(prec), device_(device), pps_(passes_per_step), max_active_(static_cast<size_t>(max_active))
{
     if (passes_per_step < 1 || passes_per_step > 64)
          throw std::runtime_error("Server(): passes_per_step must be 1..64");
     if (max_active < 1 || max_active > 4096) throw std::runtime_error("Server(): max_active must be 1..4096");
     int ndev = 0;
     check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
     if (device < 0 || device >= ndev)
          throw std::runtime_error("Server(): no HIP device " + std::to_string(device));
     DeviceScope ds(device_);
     const size_t ab = amp_bytes(prec_);
     state_cap_amps_ = state_bytes / ab;
     const uint64_t mat_entries = matrix_bytes / ab;
     if (state_cap_amps_ < 2 || mat_entries < 4) throw std::runtime_error("Server(): pools too small");
     size_t free_b = 0, total_b = 0;
     check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
     if (state_bytes + matrix_bytes + (size_t(2) << 30) > free_b)
          throw std::runtime_error("Server(): pools exceed free device memory");
     check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
     check(hipMalloc(&d_states_, state_cap_amps_ * ab), "hipMalloc(states)");
     check(hipMalloc(&d_mats_, mat_entries * ab), "hipMalloc(matrices)");
     check(hipMalloc(reinterpret_cast<void**>(&d_scratch_), kReduceScratchDoubles * sizeof(double)),
           "hipMalloc(scratch)");
     state_pool_ = RangeAlloc(state_cap_amps_);
     mat_pool_ = RangeAlloc(mat_entries);
     mat_cap_entries_ = mat_entries;
     for (auto& s: slots_) check(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming), "hipEventCreate");
     for (int i = 0; i < kPlanners; ++i) planners_.emplace_back([this] { planner_loop(); });
}

Server::~Server()
{
     {
          std::lock_guard<std::mutex> lk(qmu_);
          stop_ = true;
          for (auto& kv: reqs_)
               if (kv.second.plan) kv.second.plan->cancel = true;
     }
     qcv_.notify_all();
     for (auto& t: planners_) t.join();
     (void)hipSetDevice(device_);
     (void)hipStreamSynchronize(stream_);
     for (auto& s: slots_) {
          if (s.h_desc) (void)hipHostFree(s.h_desc);
          if (s.h_out) (void)hipHostFree(s.h_out);
          if (s.d_desc) (void)hipFree(s.d_desc);
          if (s.d_out) (void)hipFree(s.d_out);
          (void)hipEventDestroy(s.ev);
     }
     for (void* p: {d_states_, d_mats_, static_cast<void*>(d_sums_), static_cast<void*>(d_scratch_)})
          if (p) (void)hipFree(p);
     (void)hipStreamDestroy(stream_);
}

void Server::set_prog_min_qubits(int q)
{
     if (q < 0 || q > 31) throw std::runtime_error("Server: prog_min_qubits must be 0..31");
     prog_min_qubits_ = q;
}

void Server::set_heavy_qubits(int q)
{
     if (q < 1 || q > 31) throw std::runtime_error("Server: heavy_qubits must be 1..31");
     heavy_qubits_ = q;
}

void Server::set_widths(const std::vector<int>& k_by_n)
{
     for (int k: k_by_n)
          if (k < 0 || k > kMaxWidth) throw std::runtime_error("Server: fusion width must be 0..5");
     k_by_n_ = k_by_n;
}

int Server::fusion_width(int n) const
{
     const int k = (n >= 0 && static_cast<size_t>(n) < k_by_n_.size()) ? k_by_n_[n] : 0;
     return k > 0 ? k : kDefaultWidth;
}

int64_t Server::submit(int n, std::vector<BatchGate> gates, ServeKind kind, uint32_t shots,
                       uint64_t seed, std::vector<uint64_t> xmask, std::vector<uint64_t> zmask)
{
     if (n < 1 || n > kMaxServeQubits)
          throw std::runtime_error("Server: requests must have 1.." + std::to_string(kMaxServeQubits) + " qubits");
     if ((uint64_t(1) << n) > state_cap_amps_)
          throw std::runtime_error("Server: the request's state does not fit the state pool");
     if (kind == ServeKind::Samples && (shots == 0 || shots > (1u << 20)))
          throw std::runtime_error("Server: shots must be 1..2^20");
     if (kind == ServeKind::Expectation) {
          if (xmask.size() != zmask.size() || xmask.empty())
               throw std::runtime_error("Server: expectation needs one x and one z mask per term");
          const uint64_t lim = n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
          for (size_t t = 0; t < xmask.size(); ++t)
               if ((xmask[t] | zmask[t]) & ~lim) throw std::runtime_error("Server: Pauli term acts beyond the register");
     }
     // Validation and planning as for a batch circuit.
     for (size_t gi = 0; gi < gates.size(); ++gi) {
          const auto& g = gates[gi];
          const std::string where = "Server: gate " + std::to_string(gi) + ": ";
          if (g.targets.empty() || g.targets.size() + g.ctrls.size() > static_cast<size_t>(kMaxWidth))
               throw std::runtime_error(where + "needs 1..5 qubits with its controls");
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
     Req r;
     r.n = n;
     r.kind = kind;
     r.shots = shots;
     r.seed = seed;
     r.xmask = std::move(xmask);
     r.zmask = std::move(zmask);
     r.t_submit = now_s();
     const size_t ab = amp_bytes(prec_);
     std::shared_ptr<PlanState> ps;
     if (!gates.empty() && n >= std::max(prog_min_qubits_, kTileBits)) {
          ps = std::make_shared<PlanState>();
          ps->gates = std::move(gates);
          ps->n = n;
          ps->width = fusion_width(n);
          r.plan = ps;
     }
     else if (!gates.empty()) {
          std::vector<std::vector<cplx>*> mats;
          std::vector<PlannedPass> fused;
          {
               fused = plan_circuit(gates, fusion_width(n));
               for (auto& p: fused) {
                    Pass q{p.k, {0, 0, 0, 0, 0}};
                    std::memcpy(q.pos, p.pos, 5);
                    r.passes.push_back(q);
                    r.pass_mat.push_back(r.mat_entries);
                    r.mat_entries += p.m.size();
                    mats.push_back(&p.m);
               }
          }
          r.mats.resize(r.mat_entries * ab);
          uint64_t o = 0;
          for (const auto* mp: mats) {
               const auto& m = *mp;
               for (size_t i = 0; i < m.size(); ++i, ++o) {
                    if (prec_ == Prec::C128) {
                         double* d = reinterpret_cast<double*>(r.mats.data()) + 2 * o;
                         d[0] = m[i].real();
                         d[1] = m[i].imag();
                    }
                    else {
                         float* d = reinterpret_cast<float*>(r.mats.data()) + 2 * o;
                         d[0] = static_cast<float>(m[i].real());
                         d[1] = static_cast<float>(m[i].imag());
                    }
               }
          }
     }
     const int64_t id = next_id_++;
     reqs_.emplace(id, std::move(r));
     waiting_.push_back(id);
     ++stats_.submitted;
     if (ps) {
          {
               std::lock_guard<std::mutex> lk(qmu_);
               plan_queue_.push_back(ps);
          }
          qcv_.notify_one();
     }
     return id;
}

// Planner thread: plans queued tile-program requests pass by pass and
// publishes each pass (ops and matrices) under the request's lock.
void Server::planner_loop()
{
     const size_t ab = amp_bytes(prec_);
     for (;;) {
          std::shared_ptr<PlanState> ps;
          {
               std::unique_lock<std::mutex> lk(qmu_);
               qcv_.wait(lk, [&] { return stop_ || !plan_queue_.empty(); });
               if (stop_) return;
               ps = plan_queue_.front();
               plan_queue_.pop_front();
          }
          try {
               TilePassPlanner planner(gate_list(ps->gates), ps->n, kTileHigh, ps->width);
               TilePlan tp;
               while (!ps->cancel && planner.next(tp)) {
                    const PlannedTile t = fuse_tile(ps->gates, tp);
                    size_t entries = 0;
                    for (const auto& op: t.ops) entries += op.m.size();
                    std::vector<char> mb(entries * ab);
                    size_t o = 0;
                    for (const auto& op: t.ops) {
                         put_tile_matrix(mb.data() + o * ab, op.m.data(), op.k, prec_);
                         o += op.m.size();
                    }
                    std::lock_guard<std::mutex> g(ps->m);
                    TPass pass{};
                    std::memcpy(pass.hpos, t.hpos, kTileHigh);
                    pass.op_first = static_cast<uint32_t>(ps->tops.size());
                    pass.nops = static_cast<uint32_t>(t.ops.size());
                    for (const auto& op: t.ops) {
                         TileOp to{};
                         to.mat_off = ps->mat_entries;
                         to.k = static_cast<uint8_t>(op.k);
                         std::memcpy(to.rpos, op.pos, 5);
                         ps->kmax_op = std::max(ps->kmax_op, op.k);
                         ps->tops.push_back(to);
                         ps->mat_entries += op.m.size();
                    }
                    ps->mats.insert(ps->mats.end(), mb.begin(), mb.end());
                    ps->tpasses.push_back(pass);
               }
          }
          catch (const std::exception& e) {
               std::lock_guard<std::mutex> g(ps->m);
               ps->error = e.what();
          }
          std::lock_guard<std::mutex> g(ps->m);
          ps->done = true;
     }
}

void Server::ensure_slot(Slot& s, size_t desc_bytes, size_t out_bytes)
{
     if (desc_bytes > s.desc_cap) {
          const size_t cap = std::max(desc_bytes, s.desc_cap * 2);
          if (s.h_desc) check(hipHostFree(s.h_desc), "hipHostFree");
          if (s.d_desc) check(hipFree(s.d_desc), "hipFree");
          check(hipHostMalloc(reinterpret_cast<void**>(&s.h_desc), cap, hipHostMallocDefault), "hipHostMalloc");
          check(hipMalloc(reinterpret_cast<void**>(&s.d_desc), cap), "hipMalloc");
          s.desc_cap = cap;
     }
     if (out_bytes > s.out_cap) {
          const size_t cap = std::max(out_bytes, s.out_cap * 2);
          if (s.h_out) check(hipHostFree(s.h_out), "hipHostFree");
          if (s.d_out) check(hipFree(s.d_out), "hipFree");
          check(hipHostMalloc(reinterpret_cast<void**>(&s.h_out), cap, hipHostMallocDefault), "hipHostMalloc");
          check(hipMalloc(reinterpret_cast<void**>(&s.d_out), cap), "hipMalloc");
          s.out_cap = cap;
     }
}

void Server::harvest(bool wait_oldest, std::vector<int64_t>& completed)
{
     while (!inflight_.empty()) {
          Slot& s = slots_[inflight_.front()];
          if (wait_oldest) {
               check(hipEventSynchronize(s.ev), "hipEventSynchronize");
               wait_oldest = false;
          }
          else {
               const hipError_t q = hipEventQuery(s.ev);
               if (q == hipErrorNotReady) return;
               check(q, "hipEventQuery");
          }
          const double t = now_s();
          for (const auto& fo: s.finishing) {
               Req& r = reqs_.at(fo.first);
               const char* src = s.h_out + fo.second;
               if (r.kind == ServeKind::Samples) {
                    r.res_samples.resize(r.shots);
                    std::memcpy(r.res_samples.data(), src, r.shots * sizeof(uint64_t));
               }
               else if (r.kind == ServeKind::State) {
                    const uint64_t na = uint64_t(1) << r.n;
                    r.res_state.resize(na);
                    if (prec_ == Prec::C128) {
                         std::memcpy(r.res_state.data(), src, na * 16);
                    }
                    else {
                         const float* f = reinterpret_cast<const float*>(src);
                         for (uint64_t i = 0; i < na; ++i) r.res_state[i] = std::complex<double>(f[2 * i], f[2 * i + 1]);
                    }
               }
               else {
                    // <X^x Z^z> from the kernel times i^{|x & z|} (Y = i X Z)
                    static const std::complex<double> ip[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
                    r.res_exp.resize(r.xmask.size());
                    const double* d = reinterpret_cast<const double*>(src);
                    for (size_t i = 0; i < r.res_exp.size(); ++i)
                         r.res_exp[i] = ip[__builtin_popcountll(r.xmask[i] & r.zmask[i]) & 3] *
                                        std::complex<double>(d[2 * i], d[2 * i + 1]);
               }
               r.state = 3;
               r.t_done = t;
               completed.push_back(fo.first);
               ++stats_.completed;
          }
          s.finishing.clear();
          s.busy = false;
          inflight_.pop_front();
     }
}

std::vector<int64_t> Server::step(bool block)
{
     DeviceScope ds(device_);
     std::vector<int64_t> completed;
     harvest(false, completed);
     const size_t ab = amp_bytes(prec_);

     // Admission in arrival order; a request that does not fit lets later
     // ones try.
     std::vector<int64_t> admitted;
     bool heavy_active = false;
     for (int64_t id: active_)
          if (reqs_.at(id).n >= heavy_qubits_) heavy_active = true;
     for (auto it = waiting_.begin(); it != waiting_.end() && active_.size() < max_active_;) {
          Req& r = reqs_.at(*it);
          const bool heavy = r.n >= heavy_qubits_;
          if (heavy && heavy_active) {
               ++it;
               continue;
          }
          if (r.plan) {  // tile programs: admitted once the first pass is planned
               std::lock_guard<std::mutex> g(r.plan->m);
               if (!r.plan->error.empty())
                    throw std::runtime_error("Server: planning request " + std::to_string(*it) +
                                             " failed: " + r.plan->error);
               if (r.plan->tpasses.empty() && !r.plan->done) {
                    ++it;
                    continue;
               }
          }
          uint64_t so = 0, mo = 0;
          if (!state_pool_.alloc(uint64_t(1) << r.n, 256 / ab, &so)) {
               ++it;
               continue;
          }
          if (r.mat_entries && !mat_pool_.alloc(r.mat_entries, 64, &mo)) {
               state_pool_.release(so, uint64_t(1) << r.n);
               ++it;
               continue;
          }
          r.state_off = so;
          r.mat_off = mo;
          r.state = 1;
          r.t_admit = now_s();
          heavy_active = heavy_active || heavy;
          admitted.push_back(*it);
          active_.push_back(*it);
          it = waiting_.erase(it);
     }
     stats_.peak_active = std::max<uint64_t>(stats_.peak_active, active_.size());
     stats_.peak_state_bytes = std::max<uint64_t>(stats_.peak_state_bytes,
                                                  (state_cap_amps_ - state_pool_.free_total()) * ab);
     if (active_.empty()) {
          if (block && !inflight_.empty()) harvest(true, completed);
          return completed;
     }
     // Descriptors of this step.
     ItemList init;
     for (int64_t id: admitted) {
          const Req& r = reqs_.at(id);
          BatchItem it{};
          it.state_off = r.state_off;
          it.log_groups = static_cast<uint32_t>(r.n);
          init.add(it, ((uint64_t(1) << r.n) + kInitAmpsPerBlock - 1) / kInitAmpsPerBlock);
     }
     // The wave-tile form is used under the same rule as in a batch (not
     // when every target is at bit >= 20, nor for k = 5 without bit 0 (0..1
     // in complex64) among the targets, nor for states too small for a
     // block of tiles).
     std::vector<StepLayer> layers(pps_);
     int kmax = 1;
     std::vector<int64_t> finishing;
     std::vector<TileOp> step_ops;  // ops of this step's tile programs (matrix offsets into step_mats)
     std::vector<char> step_mats;   // their matrices, shipped with the step's descriptors
     uint64_t step_mat_entries = 0;
     bool work = !init.items.empty();
     for (int64_t id: active_) {
          Req& r = reqs_.at(id);
          r.adv = 0;
          if (r.plan) {
               PlanState& ps = *r.plan;
               std::lock_guard<std::mutex> g(ps.m);
               if (!ps.error.empty())
                    throw std::runtime_error("Server: planning request " + std::to_string(id) + " failed: " + ps.error);
               for (int l = 0; l < pps_ && r.next + l < ps.tpasses.size(); ++l) {
                    const TPass& tp = ps.tpasses[r.next + l];
                    if (r.state_off + (uint64_t(1) << r.n) > state_cap_amps_ || r.n < kTileBits)
                         throw std::runtime_error("Server: internal bounds check failed");
                    BatchTileProg it{};
                    it.state_off = r.state_off;
                    it.op_first = static_cast<uint32_t>(step_ops.size());
                    it.nops = static_cast<uint16_t>(tp.nops);
                    it.log_blocks = static_cast<uint8_t>(r.n - kTileBits);
                    std::memcpy(it.hpos, tp.hpos, kTileHigh);
                    for (uint32_t o = tp.op_first; o < tp.op_first + tp.nops; ++o) {
                         TileOp to = ps.tops[o];
                         const uint64_t ne = uint64_t(1) << (2 * to.k);
                         if ((to.mat_off + ne) * ab > ps.mats.size())
                              throw std::runtime_error("Server: internal bounds check failed");
                         step_mats.insert(step_mats.end(), ps.mats.begin() + to.mat_off * ab,
                                          ps.mats.begin() + (to.mat_off + ne) * ab);
                         to.mat_off = step_mat_entries;
                         step_mat_entries += ne;
                         step_ops.push_back(to);
                    }
                    layers[l].prog_kmax = std::max(layers[l].prog_kmax, ps.kmax_op);
                    layers[l].progs.add(it, uint64_t(1) << it.log_blocks);
                    ++r.adv;
                    work = true;
               }
               if (ps.done && r.next + r.adv >= ps.tpasses.size()) {
                    finishing.push_back(id);
                    work = true;
               }
               continue;
          }
          const size_t total = r.passes.size();
          for (int l = 0; l < pps_ && r.next + l < r.passes.size(); ++l) {
               const Pass& p = r.passes[r.next + l];
               const uint64_t moff = r.mat_off + r.pass_mat[r.next + l];
               const bool tile_pays = p.pos[0] < get_tile_high_target_for(p.k) &&
                                      (p.k < 5 || p.pos[0] <= (prec_ == Prec::C128 ? 0 : 1));
               if (r.state_off + (uint64_t(1) << r.n) > state_cap_amps_ ||
                   moff + (uint64_t(1) << (2 * p.k)) > mat_cap_entries_ || p.k < 1 || p.k > r.n)
                    throw std::runtime_error("Server: internal bounds check failed");
               BatchTileItem ti{};
               if (tile_pays && make_tile_item(r.n, p.k, p.pos, prec_, ti)) {
                    ti.state_off = r.state_off;
                    ti.mat_off = moff;
                    const uint64_t nb = (uint64_t(1) << (r.n - p.k)) / (64ull * batch_tile_wpb(p.k, prec_));
                    uint64_t last = (nb * batch_tile_wpb(p.k, prec_) - 1) << 6;
                    for (int q = 0; q < p.k; ++q) {
                         const uint64_t lo = last & ((uint64_t(1) << p.pos[q]) - 1);
                         last = ((last ^ lo) << 1) | lo;
                    }
                    if (nb == 0 || (uint64_t(1) << ti.log_blocks) != nb ||
                        last + 63 + ti.hoff[(1 << p.k) - 1] >= (uint64_t(1) << r.n))
                         throw std::runtime_error("Server: internal tile bounds check failed");
                    layers[l].tile[p.k].add(ti, nb);
                    continue;
               }
               BatchItem it{};
               it.state_off = r.state_off;
               it.mat_off = moff;
               it.log_groups = static_cast<uint32_t>(r.n - p.k);
               std::memcpy(it.pos, p.pos, 5);
               it.k = static_cast<uint8_t>(p.k);
               kmax = std::max(kmax, p.k);
               layers[l].direct.add(it, ((uint64_t(1) << it.log_groups) + kBatchLayerBlock - 1) / kBatchLayerBlock);
          }
          r.adv = std::min(total - std::min(total, r.next), static_cast<size_t>(pps_));
          if (r.next + pps_ >= total) finishing.push_back(id);
          work = true;
     }
     if (!work) {
          // Only tile-program requests whose next passes are still being
          // planned: nothing to launch yet.
          if (block) {
               if (!inflight_.empty()) harvest(true, completed);
               else std::this_thread::sleep_for(std::chrono::microseconds(200));
          }
          return completed;
     }
     // A free slot: wait for the oldest step when both are in flight (slots
     // are used strictly in turn, so the oldest in flight is the next one).
     if (inflight_.size() >= static_cast<size_t>(kServeSlots)) harvest(true, completed);
     const int si = next_slot_;
     next_slot_ = (next_slot_ + 1) % kServeSlots;
     Slot& slot = slots_[si];
     // Outputs of the requests that finish in this step.
     ItemList sums;
     std::vector<SampleJob> jobs;
     std::vector<std::pair<int64_t, size_t>> outs;
     size_t out_bytes = 0;
     uint64_t sums_needed = 0;
     for (int64_t id: finishing) {
          Req& r = reqs_.at(id);
          outs.push_back({id, out_bytes});
          if (r.kind == ServeKind::Samples) {
               const uint32_t lb = sample_log_blocks(r.n);
               BatchItem it{};
               it.state_off = r.state_off;
               it.mat_off = sums_needed;
               it.log_groups = static_cast<uint32_t>(r.n) - lb;
               sums.add(it, uint64_t(1) << lb);
               SampleJob j{};
               j.state_off = r.state_off;
               j.sums_off = sums_needed;
               j.out_off = out_bytes / sizeof(uint64_t);
               j.seed = r.seed;
               j.log_blocks = lb;
               j.log_block_amps = static_cast<uint32_t>(r.n) - lb;
               j.shots = r.shots;
               jobs.push_back(j);
               sums_needed += uint64_t(1) << lb;
               out_bytes += align_up(r.shots * sizeof(uint64_t), 256);
          }
          else if (r.kind == ServeKind::State) {
               out_bytes += align_up((uint64_t(1) << r.n) * ab, 256);
          }
          else {
               out_bytes += align_up(r.xmask.size() * 2 * sizeof(double), 256);
          }
     }
     if (sums_needed > sums_cap_) {
          check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
          if (d_sums_) check(hipFree(d_sums_), "hipFree");
          sums_cap_ = std::max<uint64_t>(sums_needed, 2 * sums_cap_);
          check(hipMalloc(reinterpret_cast<void**>(&d_sums_), sums_cap_ * sizeof(double)), "hipMalloc(sums)");
     }
     // Layout in the slot's staging buffer.
     const size_t o_init = 0;
     size_t o = align_up(init.items.size() * sizeof(BatchItem), 256);
     std::vector<size_t> o_layer(pps_), o_prog(pps_);
     std::vector<std::array<size_t, 6>> o_tile(pps_);
     for (int l = 0; l < pps_; ++l) {
          o_prog[l] = o;
          o += align_up(layers[l].progs.items.size() * sizeof(BatchTileProg), 256);
          o_layer[l] = o;
          o += align_up(layers[l].direct.items.size() * sizeof(BatchItem), 256);
          for (int k = 1; k <= kMaxWidth; ++k) {
               o_tile[l][k] = o;
               o += align_up(layers[l].tile[k].items.size() * sizeof(BatchTileItem), 256);
          }
     }
     const size_t o_ops = o;
     o += align_up(step_ops.size() * sizeof(TileOp), 256);
     const size_t o_tmats = o;
     o += align_up(step_mats.size(), 256);
     const size_t o_sums = o;
     o += align_up(sums.items.size() * sizeof(BatchItem), 256);
     const size_t o_jobs = o;
     o += align_up(jobs.size() * sizeof(SampleJob), 256);
     const size_t desc_bytes = std::max<size_t>(o, 256);
     ensure_slot(slot, desc_bytes, std::max<size_t>(out_bytes, 256));
     std::memcpy(slot.h_desc + o_init, init.items.data(), init.items.size() * sizeof(BatchItem));
     std::memcpy(slot.h_desc + o_ops, step_ops.data(), step_ops.size() * sizeof(TileOp));
     std::memcpy(slot.h_desc + o_tmats, step_mats.data(), step_mats.size());
     for (int l = 0; l < pps_; ++l) {
          std::memcpy(slot.h_desc + o_prog[l], layers[l].progs.items.data(),
                      layers[l].progs.items.size() * sizeof(BatchTileProg));
          std::memcpy(slot.h_desc + o_layer[l], layers[l].direct.items.data(),
                      layers[l].direct.items.size() * sizeof(BatchItem));
          for (int k = 1; k <= kMaxWidth; ++k)
               std::memcpy(slot.h_desc + o_tile[l][k], layers[l].tile[k].items.data(),
                           layers[l].tile[k].items.size() * sizeof(BatchTileItem));
     }
     std::memcpy(slot.h_desc + o_sums, sums.items.data(), sums.items.size() * sizeof(BatchItem));
     std::memcpy(slot.h_desc + o_jobs, jobs.data(), jobs.size() * sizeof(SampleJob));

     // Matrices of the newly admitted requests, then the step's launches.
     for (int64_t id: admitted) {
          const Req& r = reqs_.at(id);
          if (r.mat_entries)
               check(hipMemcpyAsync(static_cast<char*>(d_mats_) + r.mat_off * ab, r.mats.data(),
                                    r.mat_entries * ab, hipMemcpyHostToDevice, stream_),
                     "hipMemcpyAsync(matrices)");
     }
     check(hipMemcpyAsync(slot.d_desc, slot.h_desc, desc_bytes, hipMemcpyHostToDevice, stream_),
           "hipMemcpyAsync(descriptors)");
     auto d_items = [&](size_t off) { return reinterpret_cast<const BatchItem*>(slot.d_desc + off); };
     if (!init.items.empty()) {
          check(launch_init_states(d_states_, prec_, d_items(o_init), static_cast<int>(init.items.size()),
                                   init.blocks, stream_),
                "launch_init_states");
          ++stats_.launches;
     }
     for (int l = 0; l < pps_; ++l) {
          if (!layers[l].progs.items.empty()) {
               check(launch_batch_tprog(d_states_, prec_, slot.d_desc + o_tmats,
                                        reinterpret_cast<const BatchTileProg*>(slot.d_desc + o_prog[l]),
                                        static_cast<int>(layers[l].progs.items.size()),
                                        reinterpret_cast<const TileOp*>(slot.d_desc + o_ops),
                                        layers[l].prog_kmax, layers[l].progs.blocks, stream_),
                     "launch_batch_tprog");
               ++stats_.launches;
          }
          for (int k = 1; k <= kMaxWidth; ++k) {
               const TileList& tl = layers[l].tile[k];
               if (tl.items.empty()) continue;
               check(launch_batch_tile(d_states_, prec_, d_mats_,
                                       reinterpret_cast<const BatchTileItem*>(slot.d_desc + o_tile[l][k]),
                                       static_cast<int>(tl.items.size()), k, tl.blocks, stream_),
                     "launch_batch_tile");
               ++stats_.launches;
          }
          const ItemList& dl = layers[l].direct;
          if (dl.items.empty()) continue;
          check(launch_serve_layer(d_states_, prec_, d_mats_, d_items(o_layer[l]),
                                   static_cast<int>(dl.items.size()), dl.blocks, kmax, stream_),
                "launch_serve_layer");
          ++stats_.launches;
     }
     if (!jobs.empty()) {
          check(launch_block_sums_multi(d_states_, prec_, d_items(o_sums), static_cast<int>(sums.items.size()),
                                        sums.blocks, d_sums_, stream_),
                "launch_block_sums_multi");
          check(launch_sample(d_states_, prec_, reinterpret_cast<const SampleJob*>(slot.d_desc + o_jobs),
                              static_cast<int>(jobs.size()), d_sums_,
                              reinterpret_cast<uint64_t*>(slot.d_out), stream_),
                "launch_sample");
          stats_.launches += 2;
     }
     for (const auto& fo: outs) {
          const Req& r = reqs_.at(fo.first);
          if (r.kind != ServeKind::Expectation) continue;
          const void* psi = static_cast<const char*>(d_states_) + r.state_off * ab;
          double* out = reinterpret_cast<double*>(slot.d_out + fo.second);
          for (size_t t = 0; t < r.xmask.size(); ++t) {
               check(launch_pauli_expectation(psi, prec_, uint64_t(1) << r.n, r.xmask[t], r.zmask[t],
                                              d_scratch_, out + 2 * t, stream_),
                     "launch_pauli_expectation");
               ++stats_.launches;
          }
     }
     if (out_bytes)
          check(hipMemcpyAsync(slot.h_out, slot.d_out, out_bytes, hipMemcpyDeviceToHost, stream_),
                "hipMemcpyAsync(outputs)");
     // States go straight from the pool to their place in the pinned output
     // buffer, after the copy above (same stream).
     for (const auto& fo: outs) {
          const Req& r = reqs_.at(fo.first);
          if (r.kind != ServeKind::State) continue;
          const uint64_t bytes = (uint64_t(1) << r.n) * ab;
          if (fo.second + bytes > slot.out_cap) throw std::runtime_error("Server: internal output bounds check failed");
          check(hipMemcpyAsync(slot.h_out + fo.second, static_cast<const char*>(d_states_) + r.state_off * ab, bytes,
                               hipMemcpyDeviceToHost, stream_),
                "hipMemcpyAsync(state)");
     }
     check(hipEventRecord(slot.ev, stream_), "hipEventRecord");
     slot.busy = true;
     slot.finishing = outs;
     inflight_.push_back(si);
     ++stats_.steps;

     // Progress; finished requests give their memory back (later launches on
     // the stream run after this step's outputs).
     std::vector<int64_t> still;
     for (int64_t id: active_) {
          Req& r = reqs_.at(id);
          r.next += r.adv;
          if (std::find(finishing.begin(), finishing.end(), id) != finishing.end()) {
               r.state = 2;
               state_pool_.release(r.state_off, uint64_t(1) << r.n);
               if (r.mat_entries) mat_pool_.release(r.mat_off, r.mat_entries);
               std::vector<char>().swap(r.mats);
               r.plan.reset();
          }
          else {
               still.push_back(id);
          }
     }
     active_.swap(still);
     if (block && active_.empty() && waiting_.empty()) harvest(true, completed);
     return completed;
}

bool Server::done(int64_t id) const
{
     auto it = reqs_.find(id);
     if (it == reqs_.end()) throw std::runtime_error("Server: unknown request " + std::to_string(id));
     return it->second.state == 3;
}

const std::vector<uint64_t>& Server::samples(int64_t id) const
{
     if (!done(id)) throw std::runtime_error("Server: request " + std::to_string(id) + " is not done");
     const Req& r = reqs_.at(id);
     if (r.kind != ServeKind::Samples) throw std::runtime_error("Server: request has no samples");
     return r.res_samples;
}

const std::vector<std::complex<double>>& Server::expectations(int64_t id) const
{
     if (!done(id)) throw std::runtime_error("Server: request " + std::to_string(id) + " is not done");
     const Req& r = reqs_.at(id);
     if (r.kind != ServeKind::Expectation) throw std::runtime_error("Server: request has no expectations");
     return r.res_exp;
}

const std::vector<std::complex<double>>& Server::state(int64_t id) const
{
     if (!done(id)) throw std::runtime_error("Server: request " + std::to_string(id) + " is not done");
     const Req& r = reqs_.at(id);
     if (r.kind != ServeKind::State) throw std::runtime_error("Server: request has no state output");
     return r.res_state;
}

void Server::times(int64_t id, double* t_submit, double* t_admit, double* t_done) const
{
     auto it = reqs_.find(id);
     if (it == reqs_.end()) throw std::runtime_error("Server: unknown request " + std::to_string(id));
     *t_submit = it->second.t_submit;
     *t_admit = it->second.t_admit;
     *t_done = it->second.t_done;
}

void Server::forget(int64_t id)
{
     auto it = reqs_.find(id);
     if (it == reqs_.end()) return;
     if (it->second.state != 3) throw std::runtime_error("Server: request is still running");
     reqs_.erase(it);
}

}  // namespace hq
