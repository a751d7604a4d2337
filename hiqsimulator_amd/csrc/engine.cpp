// Host side of the single-GPU state-vector engine.  See engine.hpp.
#include "engine.hpp"

#include "batch.hpp"
#include "sched.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace hq {
namespace {

constexpr double kMaxFloatError = 1e-12;  // SimulatorMPI.cpp:75
constexpr uint64_t kMaxBlocks = uint64_t(1) << 15;  // SimulatorMPI.cpp:903
constexpr int kMaxOuts = 256;                        // batched reduction outputs

void check(hipError_t e, const char* what)
{
     if (e != hipSuccess) {
          std::ostringstream os;
          os << what << ": " << hipGetErrorString(e);
          throw std::runtime_error(os.str());
     }
}

std::string id_list(const std::vector<int64_t>& v)
{
     std::ostringstream os;
     os << "[";
     for (size_t i = 0; i < v.size(); ++i) os << (i ? ", " : "") << v[i];
     os << "]";
     return os.str();
}

struct DeviceScope {
     explicit DeviceScope(int dev) { check(hipSetDevice(dev), "hipSetDevice"); }
};

}  // namespace

Engine::Engine(uint64_t seed, int64_t max_local, int64_t max_cluster, Prec prec, int device)
    : prec_(prec),
      device_(device),
      max_local_(static_cast<size_t>(std::max<int64_t>(0, max_local))),
      max_cluster_(static_cast<size_t>(std::max<int64_t>(0, max_cluster))),
      min_local_(static_cast<size_t>(std::max<int64_t>(0, max_cluster))),
      // std::mt19937 keeps the seed modulo 2^32 (SimulatorMPI.cpp:88).
      rng_(static_cast<std::mt19937::result_type>(seed))
{
     if (max_local < 0 || max_cluster < 0)
          throw std::runtime_error("SimulatorMPI(): max_local and max_cluster_size must be >= 0");
     int ndev = 0;
     check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
     if (device < 0 || device >= ndev)
          throw std::runtime_error("SimulatorMPI(): no HIP device " + std::to_string(device));
     DeviceScope ds(device_);
     check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
     check(hipHostMalloc(reinterpret_cast<void**>(&h_ring_), kRing * kSlotBytes,
                         hipHostMallocDefault),
           "hipHostMalloc(ring)");
     check(hipMalloc(reinterpret_cast<void**>(&d_ring_), kRing * kSlotBytes), "hipMalloc(ring)");
     for (int i = 0; i < kRing; ++i)
          check(hipEventCreateWithFlags(&ring_ev_[i], hipEventDisableTiming), "hipEventCreate");
     check(hipMalloc(reinterpret_cast<void**>(&d_scratch_),
                     (kReduceScratchDoubles + 2 * kMaxOuts) * sizeof(double)),
           "hipMalloc(scratch)");
     check(hipHostMalloc(reinterpret_cast<void**>(&h_out_), 2 * kMaxOuts * sizeof(double),
                         hipHostMallocDefault),
           "hipHostMalloc(out)");
     check(hipMalloc(reinterpret_cast<void**>(&d_blocks_), kMaxBlocks * sizeof(double)),
           "hipMalloc(blocks)");
     check(hipHostMalloc(reinterpret_cast<void**>(&h_blocks_), kMaxBlocks * sizeof(double),
                         hipHostMallocDefault),
           "hipHostMalloc(blocks)");
     // The empty register is the one-amplitude state [1] (SimulatorMPI.cpp:96-98).
     ensure_state(1);
     check(launch_fill(d_state_, prec_, 1, 1.0, 0.0, stream_), "fill");
     n_amps_ = 1;
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

Engine::~Engine()
{
     (void)hipSetDevice(device_);
     if (stream_) (void)hipStreamSynchronize(stream_);
     for (auto& p: pending_) {
          (void)hipEventDestroy(p.start);
          (void)hipEventDestroy(p.stop);
     }
     for (auto e: free_events_) (void)hipEventDestroy(e);
     for (int i = 0; i < kRing; ++i)
          if (ring_ev_[i]) (void)hipEventDestroy(ring_ev_[i]);
     if (d_state_) (void)hipFree(d_state_);
     if (d_tmp_) (void)hipFree(d_tmp_);
     for (void* p: retired_) (void)hipFree(p);
     if (d_snap_) (void)hipFree(d_snap_);
     if (d_ring_) (void)hipFree(d_ring_);
     if (h_ring_) (void)hipHostFree(h_ring_);
     if (d_scratch_) (void)hipFree(d_scratch_);
     if (h_out_) (void)hipHostFree(h_out_);
     if (d_blocks_) (void)hipFree(d_blocks_);
     if (h_blocks_) (void)hipHostFree(h_blocks_);
     if (h_stage_) (void)hipHostFree(h_stage_);
     if (d_samples_) (void)hipFree(d_samples_);
     for (auto& ps: prog_) {
          if (ps.ev) (void)hipEventDestroy(ps.ev);
          if (ps.h) (void)hipHostFree(ps.h);
          if (ps.d) (void)hipFree(ps.d);
     }
     if (stream_) (void)hipStreamDestroy(stream_);
}

// ---------------------------------------------------------------------------
// Buffers
// ---------------------------------------------------------------------------
// A state buffer that may be viewed is retired instead of freed while a pin
// is held; collect_retired() frees the retired buffers once no pin is left.
void Engine::release_buffer(void*& p)
{
     if (!p) return;
     if (view_pins_.load() > 0) {
          retired_.push_back(p);
     }
     else {
          check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
          check(hipFree(p), "hipFree");
     }
     p = nullptr;
}

void Engine::collect_retired()
{
     if (retired_.empty() || view_pins_.load() > 0) return;
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     for (void* p: retired_) check(hipFree(p), "hipFree");
     retired_.clear();
}

void Engine::ensure_state(uint64_t n)
{
     collect_retired();
     if (n <= cap_state_) return;
     void* p = nullptr;
     check(hipMalloc(&p, n * amp_bytes(prec_)), "hipMalloc(state)");
     if (d_state_) {
          check(launch_copy(p, d_state_, prec_, n_amps_, stream_), "copy");
          release_buffer(d_state_);
     }
     d_state_ = p;
     cap_state_ = n;
}

void Engine::ensure_tmp(uint64_t n)
{
     collect_retired();
     if (n <= cap_tmp_) return;
     release_buffer(d_tmp_);
     cap_tmp_ = 0;
     check(hipMalloc(&d_tmp_, n * amp_bytes(prec_)), "hipMalloc(tmp)");
     cap_tmp_ = n;
}

void Engine::swap_tmp()
{
     std::swap(d_state_, d_tmp_);
     std::swap(cap_state_, cap_tmp_);
}

void Engine::commit_tmp()
{
     if (view_pins_.load() > 0) {
          begin_timed(kFillCopy);
          check(launch_copy(d_state_, d_tmp_, prec_, n_amps_, stream_), "copy");
          end_timed();
     }
     else {
          swap_tmp();
     }
}

const void* Engine::upload_matrix(const Mat& m, const void* desc, size_t desc_bytes)
{
     if (desc_bytes > kSlotBytes - kMatBytes) throw std::runtime_error("upload_matrix(): descriptor too large");
     const int slot = ring_next_;
     ring_next_ = (ring_next_ + 1) % kRing;
     if (ring_used_[slot]) check(hipEventSynchronize(ring_ev_[slot]), "hipEventSynchronize");
     char* h = h_ring_ + slot * kSlotBytes;
     const size_t n = m.dim * m.dim;
     size_t bytes;
     if (prec_ == Prec::C128) {
          auto* o = reinterpret_cast<double*>(h);
          for (size_t i = 0; i < n; ++i) {
               o[2 * i] = m.a[i].real();
               o[2 * i + 1] = m.a[i].imag();
          }
          bytes = n * 16;
     }
     else {
          auto* o = reinterpret_cast<float*>(h);
          for (size_t i = 0; i < n; ++i) {
               o[2 * i] = static_cast<float>(m.a[i].real());
               o[2 * i + 1] = static_cast<float>(m.a[i].imag());
          }
          bytes = n * 8;
     }
     if (desc_bytes) {
          std::memcpy(h + kMatBytes, desc, desc_bytes);
          bytes = kMatBytes + desc_bytes;
     }
     char* d = d_ring_ + slot * kSlotBytes;
     check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync(matrix)");
     check(hipEventRecord(ring_ev_[slot], stream_), "hipEventRecord");
     ring_used_[slot] = true;
     return d;
}

void Engine::read_out(double* host, size_t count)
{
     check(hipMemcpyAsync(host, d_scratch_ + kReduceScratchDoubles, count * sizeof(double),
                          hipMemcpyDeviceToHost, stream_),
           "hipMemcpyAsync(out)");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------
// Timing
// ---------------------------------------------------------------------------
void Engine::begin_timed(int cls)
{
     if (!timing_) return;
     hipEvent_t ev[2];
     for (auto& e: ev) {
          if (!free_events_.empty()) {
               e = free_events_.back();
               free_events_.pop_back();
          }
          else {
               check(hipEventCreate(&e), "hipEventCreate");
          }
     }
     check(hipEventRecord(ev[0], stream_), "hipEventRecord");
     pending_.push_back(TimedLaunch{cls, ev[0], ev[1]});
}

void Engine::end_timed()
{
     if (!timing_) return;
     check(hipEventRecord(pending_.back().stop, stream_), "hipEventRecord");
}

void Engine::set_timing(bool on)
{
     DeviceScope ds(device_);
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     timing_ = on;
     for (int c = 0; c < kNumKernelClasses; ++c) {
          t_count_[c] = 0;
          t_ms_[c] = 0;
     }
     for (auto& p: pending_) {
          free_events_.push_back(p.start);
          free_events_.push_back(p.stop);
     }
     pending_.clear();
}

void Engine::kernel_time(int cls, uint64_t& count, double& total_ms)
{
     DeviceScope ds(device_);
     if (cls < 0 || cls >= kNumKernelClasses) throw std::runtime_error("kernel_time(): bad class");
     settle();
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     for (auto& p: pending_) {
          float ms = 0;
          check(hipEventElapsedTime(&ms, p.start, p.stop), "hipEventElapsedTime");
          t_count_[p.cls] += 1;
          t_ms_[p.cls] += ms;
          free_events_.push_back(p.start);
          free_events_.push_back(p.stop);
     }
     pending_.clear();
     count = t_count_[cls];
     total_ms = t_ms_[cls];
}

void Engine::synchronize()
{
     DeviceScope ds(device_);
     settle();
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------
// Qubit bookkeeping
// ---------------------------------------------------------------------------
size_t Engine::find(int64_t id) const
{
     auto it = std::find(locals_.begin(), locals_.end(), id);
     return it == locals_.end() ? static_cast<size_t>(-1) : static_cast<size_t>(it - locals_.begin());
}

size_t Engine::find_sure(int64_t id) const
{
     size_t p = find(id);
     if (p == static_cast<size_t>(-1))
          throw std::runtime_error("ArrayFindSure(): Can't find " + std::to_string(id) + " in "
                                   + id_list(locals_));
     return p;
}

void Engine::allocate_qubit(int64_t id) { allocate_qureg({id}, cplx(0.0, 0.0)); }

void Engine::allocate_qureg(const std::vector<int64_t>& ids, cplx init)
{
     DeviceScope ds(device_);
     if (init != 0.0 && !locals_.empty())
          throw std::runtime_error(
              "AllocateQureg(): initialization of only first qureg is supported");
     // A qubit can be allocated while fewer than max(min_local, max_local)
     // are present (SimulatorMPI.cpp:190-212 with no global slots).
     const size_t limit = std::max(min_local_, max_local_);
     const size_t room = locals_.size() < limit ? limit - locals_.size() : 0;
     const size_t count = std::min(room, ids.size());
     settle();
     if (count > 0) {
          if (locals_.size() + count > 62)
               throw std::runtime_error("AllocateQubit(): state vector too large");
          const uint64_t n_new = n_amps_ << count;
          ensure_state(n_new);
          begin_timed(kFillCopy);
          check(launch_fill(static_cast<char*>(d_state_) + n_amps_ * amp_bytes(prec_), prec_,
                            n_new - n_amps_, 0.0, 0.0, stream_),
                "fill");
          end_timed();
          n_amps_ = n_new;
          locals_.insert(locals_.end(), ids.begin(), ids.begin() + static_cast<long>(count));
     }
     if (count < ids.size())
          throw std::runtime_error("AllocateQubit(): can't allocate more than "
                                   + std::to_string(locals_.size()) + " qubits");
     if (init != 0.0) {
          begin_timed(kFillCopy);
          check(launch_fill(d_state_, prec_, n_amps_, init.real(), init.imag(), stream_), "fill");
          end_timed();
     }
}

void Engine::deallocate_qubit(int64_t id)
{
     DeviceScope ds(device_);
     const size_t pos = find_sure(id);
     settle();
     begin_timed(kReduce);
     check(launch_two_bucket(d_state_, prec_, n_qubits(), static_cast<int>(pos), d_scratch_,
                             d_scratch_ + kReduceScratchDoubles, stream_),
           "two_bucket");
     end_timed();
     read_out(h_out_, 2);
     const double s0 = h_out_[0], s1 = h_out_[1];
     if (!((s0 > kMaxFloatError) ^ (s1 > kMaxFloatError)))
          throw std::runtime_error("DeallocateLocalQubit(): qubit " + std::to_string(id)
                                   + " is entangled");
     const int keep = s0 > kMaxFloatError ? 0 : 1;
     ensure_tmp(n_amps_ / 2);
     begin_timed(kRemoveBit);
     check(launch_remove_bit(d_tmp_, d_state_, prec_, n_amps_, static_cast<int>(pos), keep,
                             stream_),
           "remove_bit");
     end_timed();
     swap_tmp();
     n_amps_ /= 2;
     locals_.erase(locals_.begin() + static_cast<long>(pos));
}

void Engine::set_qubits_perm(const std::vector<int64_t>& ids)
{
     if (ids.size() != locals_.size())
          throw std::runtime_error("SetQubitsPermutation(): expected "
                                   + std::to_string(locals_.size()) + " ids");
     locals_ = ids;
}

void Engine::swap_qubits(const std::vector<int64_t>& pairs)
{
     // With one device every qubit is local: any (global, local) pair names a
     // qubit that is not global (SimulatorMPI.cpp:1092-1096).
     if (pairs.empty()) return;
     throw std::runtime_error("ArrayFindSure(): Can't find " + std::to_string(pairs[0])
                              + " in [] (no global qubits on a single device)");
}

void Engine::permute_qubits(const std::vector<int64_t>& ids)
{
     DeviceScope ds(device_);
     if (ids.size() != locals_.size())
          throw std::runtime_error("permute_qubits(): expected " + std::to_string(locals_.size())
                                   + " ids");
     std::vector<uint8_t> src(ids.size());
     uint64_t seen = 0;
     bool identity = true;
     for (size_t p = 0; p < ids.size(); ++p) {
          const size_t s = find_sure(ids[p]);
          src[p] = static_cast<uint8_t>(s);
          seen |= uint64_t(1) << s;
          identity = identity && s == p;
     }
     if (seen != (ids.empty() ? 0 : ((uint64_t(1) << ids.size()) - 1)))
          throw std::runtime_error("permute_qubits(): ids must be a permutation of the qubits");
     if (identity) return;
     settle();
     ensure_tmp(n_amps_);
     begin_timed(kPermute);
     check(launch_permute(d_tmp_, d_state_, prec_, n_qubits(), src.data(), stream_), "permute");
     end_timed();
     commit_tmp();
     locals_ = ids;
}

// ---------------------------------------------------------------------------
// Gates
// ---------------------------------------------------------------------------
void Engine::apply_gate(Mat m, std::vector<int64_t> ids, const std::vector<int64_t>& ctrls)
{
     if (ids.size() > 62 || m.dim != (size_t(1) << ids.size()))
          throw std::runtime_error("apply_controlled_gate(): a " + std::to_string(m.dim) + "x"
                                   + std::to_string(m.dim) + " matrix cannot act on "
                                   + std::to_string(ids.size()) + " qubits");
     const bool diag = is_diagonal(m);
     std::vector<int64_t> lc;
     for (auto c: ctrls)
          if (find(c) != static_cast<size_t>(-1)) lc.push_back(c);
     if (ids.size() + lc.size() > max_cluster_) {
          // A gate wider than a cluster runs alone with a control mask
          // (SimulatorMPI.cpp:730-756).
          run();
          fusion_.insert(std::move(m), diag, std::move(ids), ctrls);
          return;
     }
     add_controls(m, ids, lc);
     fusion_.insert(std::move(m), diag, std::move(ids), {});
}

void Engine::run()
{
     DeviceScope ds(device_);
     // A pending fusion wider than a pass can apply is rejected before its
     // 2^n x 2^n matrix is built (the reference builds it first and fails
     // with the same message; at 15 qubits that is 16 GiB of host work).
     if (fusion_.num_qubits() > 5)
          throw std::runtime_error("Run(): cannot apply " + std::to_string(fusion_.num_qubits())
                                   + " qubits gate");
     if (pass_mode_ == kPassDeferred && fusion_.num_items() > 0 && n_qubits() >= kTileBits) {
          // The cluster's gates are queued unfused, each with the shared
          // controls: the tile planner then packs gates (not whole clusters)
          // into the tiles, as apply_tiled does.
          uint64_t cmask = 0;
          for (auto c: fusion_.shared_controls()) {
               const size_t p = find(c);
               if (p != static_cast<size_t>(-1)) cmask |= uint64_t(1) << p;
          }
          if (fusion_.num_qubits() + static_cast<size_t>(__builtin_popcountll(cmask)) <= 5) {
               std::vector<int64_t> cpos;
               for (int p = 0; p < 64; ++p)
                    if ((cmask >> p) & 1) cpos.push_back(p);
               std::vector<Mat> ms;
               std::vector<std::vector<int64_t>> ids;
               fusion_.take_items(ms, ids);
               for (size_t i = 0; i < ms.size(); ++i) {
                    PosGate g;
                    g.m = std::move(ms[i]);
                    for (auto id: ids[i]) g.t.push_back(static_cast<int64_t>(find_sure(id)));
                    g.c = cpos;
                    queue_.push_back(std::move(g));
               }
               if (queue_.size() >= kQueueHigh) drain_queue(kQueueLow);
               return;
          }
     }
     Mat m;
     std::vector<int64_t> ids, ctrls;
     bool diag = false;
     fusion_.perform(m, ids, ctrls, diag);
     uint64_t cmask = 0;
     for (auto c: ctrls) {
          const size_t p = find(c);
          if (p != static_cast<size_t>(-1)) cmask |= uint64_t(1) << p;
     }
     std::vector<uint32_t> pos(ids.size());
     for (size_t i = 0; i < ids.size(); ++i) pos[i] = static_cast<uint32_t>(find_sure(ids[i]));
     if (ids.size() > 5)
          throw std::runtime_error("Run(): cannot apply " + std::to_string(ids.size())
                                   + " qubits gate");
     if (ids.empty() && m(0, 0) == cplx(1.0, 0.0)) {
          fusion_ = Fusion();
          return;
     }
     settle();
     if (ids.empty()) {
          begin_timed(kScale);
          check(launch_scale(d_state_, prec_, n_amps_, m(0, 0).real(), m(0, 0).imag(), stream_), "scale");
          end_timed();
     }
     else {
          dense_or_diag(m, pos, cmask, diag);
     }
     fusion_ = Fusion();
}

void Engine::set_pass_mode(int mode)
{
     if (mode != kPassEach && mode != kPassDeferred)
          throw std::runtime_error("set_pass_mode(): mode must be 0 (one pass per run) or 1 (deferred tile passes)");
     if (mode != pass_mode_) {
          DeviceScope ds(device_);
          settle();
     }
     pass_mode_ = mode;
}

void Engine::drain_queue(size_t keep)
{
     if (queue_.size() <= keep || queue_.empty()) return;
     int width = 4;
     for (const auto& g: queue_) width = std::max(width, static_cast<int>(g.t.size() + g.c.size()));
     tile_items(queue_, width, keep);
}

void Engine::dense_or_diag(const Mat& m, const std::vector<uint32_t>& pos, uint64_t ctrlmask,
                           bool diag)
{
     DenseDesc d{};
     d.k = static_cast<int>(pos.size());
     uint64_t all = ctrlmask;
     for (size_t l = 0; l < pos.size(); ++l) {
          d.pos[l] = pos[l];
          all |= uint64_t(1) << pos[l];
     }
     d.ctrlmask = ctrlmask;
     for (int p = 0; p < 64; ++p)
          if ((all >> p) & 1) {
               if (d.ins.n >= 63) throw std::runtime_error("Run(): too many controls");
               d.ins.p[d.ins.n++] = static_cast<uint8_t>(p);
          }
     if (diag) {
          double dv[64];
          for (size_t i = 0; i < m.dim; ++i) {
               dv[2 * i] = m(i, i).real();
               dv[2 * i + 1] = m(i, i).imag();
          }
          begin_timed(kDiag);
          check(launch_diag(d_state_, prec_, n_qubits(), dv, d, stream_), "diag pass");
          end_timed();
     }
     else if (!dense5_as_tile(m, pos, ctrlmask)) {
          const void* dm = upload_matrix(m);
          begin_timed(kDense1 + d.k - 1);
          check(launch_dense(d_state_, prec_, n_qubits(), dm, d, stream_), "dense pass");
          end_timed();
     }
}

// A 5-qubit pass without controls, on 12 or more qubits, runs as a tile pass
// with one op: the tile holds bits 0..5 plus the targets at or above bit 6
// (filled up with the lowest other bits), and the op runs on the matrix cores
// as three real products.  The dense forms keep the 32-amplitude group in
// registers on the vector units; over random placements on a 30-qubit state
// the tile pass takes 6.8 against 8.3 ms (complex128) and 3.2 against 4.4 ms
// (complex64).  Timed as a 5-qubit dense pass.
bool Engine::dense5_as_tile(const Mat& m, const std::vector<uint32_t>& pos, uint64_t ctrlmask)
{
     const int n = n_qubits();
     if (pos.size() != 5 || ctrlmask != 0 || n < kTileBits || !get_dense5_tile())
          return false;
     // matrix bit l' <-> the l'-th lowest position (the tile op wants ascending ranks)
     int order[5] = {0, 1, 2, 3, 4};
     std::sort(order, order + 5, [&](int a, int b) { return pos[a] < pos[b]; });
     // (entries in the tile kernels' matrix layout, tile_mat_slot)
     Mat ms(32);
     auto old_index = [&](size_t i) {
          size_t o = 0;
          for (int l = 0; l < 5; ++l)
               if ((i >> l) & 1) o |= size_t(1) << order[l];
          return o;
     };
     for (size_t i = 0; i < 32; ++i)
          for (size_t j = 0; j < 32; ++j)
               ms.a[tile_mat_slot(5, static_cast<uint32_t>(i * 32 + j))] = m(old_index(i), old_index(j));
     uint64_t high = 0;
     for (auto p: pos)
          if (p >= 6) high |= uint64_t(1) << p;
     for (int p = 6; p < n && __builtin_popcountll(high) < kTileHigh; ++p) high |= uint64_t(1) << p;
     TileHigh th{};
     int rank[64];
     for (int p = 0; p < 64; ++p) rank[p] = p < 6 ? p : -1;
     int jh = 0;
     for (int p = 6; p < n; ++p)
          if ((high >> p) & 1) {
               th.hpos[jh] = static_cast<uint8_t>(p);
               rank[p] = 6 + jh;
               ++jh;
          }
     if (jh != kTileHigh) throw std::runtime_error("Run(): internal tile layout error");
     TileOp op{};
     op.mat_off = 0;
     op.k = 5;
     for (int l = 0; l < 5; ++l) {
          const int r = rank[pos[order[l]]];
          if (r < 0) throw std::runtime_error("Run(): internal tile layout error");
          op.rpos[l] = static_cast<uint8_t>(r);
     }
     const char* dm = static_cast<const char*>(upload_matrix(ms, &op, sizeof(TileOp)));
     begin_timed(kDense5);
     check(launch_tile_pass(d_state_, prec_, n, th, reinterpret_cast<const TileOp*>(dm + kMatBytes), 1, 5, dm,
                            stream_),
           "dense pass (tile)");
     end_timed();
     return true;
}

void Engine::apply_matrix(const Mat& m, const std::vector<uint32_t>& pos, uint64_t ctrlmask,
                          bool diag)
{
     DeviceScope ds(device_);
     const size_t k = pos.size();
     if (k < 1 || k > 5 || m.dim != (size_t(1) << k))
          throw std::runtime_error("apply_matrix(): need 1..5 positions and a matching matrix");
     uint64_t tm = 0;
     for (auto p: pos) {
          if (p >= locals_.size() || ((tm >> p) & 1))
               throw std::runtime_error("apply_matrix(): bad or repeated bit position");
          tm |= uint64_t(1) << p;
     }
     if ((ctrlmask & tm) || (locals_.size() < 64 && (ctrlmask >> locals_.size())))
          throw std::runtime_error("apply_matrix(): bad control mask");
     settle();
     dense_or_diag(m, pos, ctrlmask, diag && is_diagonal(m));
}

// ---------------------------------------------------------------------------
// Measurement, probabilities, amplitudes
// ---------------------------------------------------------------------------
void Engine::mask_of(const std::vector<int64_t>& ids, const std::vector<uint8_t>& bits,
                     uint64_t& msk, uint64_t& val) const
{
     msk = val = 0;
     for (size_t i = 0; i < ids.size(); ++i) {
          const size_t p = find_sure(ids[i]);
          msk |= uint64_t(1) << p;
          if (bits[i]) val |= uint64_t(1) << p;
     }
}

double Engine::masked_norm(uint64_t msk, uint64_t val)
{
     begin_timed(kReduce);
     check(launch_masked_norm(d_state_, prec_, n_qubits(), msk, val, d_scratch_,
                              d_scratch_ + kReduceScratchDoubles, stream_),
           "masked_norm");
     end_timed();
     read_out(h_out_, 1);
     return h_out_[0];
}

void Engine::normalize(double norm, uint64_t msk, uint64_t val)
{
     const double f = 1.0 / std::sqrt(norm);
     begin_timed(kNormalize);
     check(launch_normalize(d_state_, prec_, n_amps_, msk, val, f, stream_), "normalize");
     end_timed();
}

std::vector<uint8_t> Engine::measure(const std::vector<int64_t>& ids)
{
     DeviceScope ds(device_);
     settle();
     const uint64_t nb = std::min<uint64_t>(n_amps_, kMaxBlocks);
     const uint64_t K = n_amps_ / nb;
     begin_timed(kReduce);
     check(launch_block_sums(d_state_, prec_, n_amps_, nb, d_blocks_, stream_), "block_sums");
     end_timed();
     check(hipMemcpyAsync(h_blocks_, d_blocks_, nb * sizeof(double), hipMemcpyDeviceToHost,
                          stream_),
           "hipMemcpyAsync(blocks)");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     std::partial_sum(h_blocks_, h_blocks_ + nb, h_blocks_);
     const double u = dist_(rng_);
     uint64_t b = 0;
     for (; b < nb; ++b)
          if (u <= h_blocks_[b]) break;
     if (b == nb) {
          std::ostringstream os;
          os << "MeasureQubits(): Random number " << u << " > norm partial sum "
             << h_blocks_[nb - 1];
          throw std::runtime_error(os.str());
     }
     // In-block scan in index order, starting from the previous prefix
     // (SimulatorMPI.cpp:943-961).
     const size_t ab = amp_bytes(prec_);
     if (K * ab > stage_bytes_) {
          if (h_stage_) check(hipHostFree(h_stage_), "hipHostFree");
          h_stage_ = nullptr;
          check(hipHostMalloc(reinterpret_cast<void**>(&h_stage_), K * ab, hipHostMallocDefault),
                "hipHostMalloc(stage)");
          stage_bytes_ = K * ab;
     }
     check(hipMemcpyAsync(h_stage_, static_cast<char*>(d_state_) + b * K * ab, K * ab,
                          hipMemcpyDeviceToHost, stream_),
           "hipMemcpyAsync(block)");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     double shift = b > 0 ? h_blocks_[b - 1] : 0.0;
     uint64_t k = b * K;
     for (uint64_t j = 0; j < K; ++j, ++k) {
          double re, im;
          if (prec_ == Prec::C128) {
               re = reinterpret_cast<const double*>(h_stage_)[2 * j];
               im = reinterpret_cast<const double*>(h_stage_)[2 * j + 1];
          }
          else {
               re = reinterpret_cast<const float*>(h_stage_)[2 * j];
               im = reinterpret_cast<const float*>(h_stage_)[2 * j + 1];
          }
          shift += re * re + im * im;
          if (shift >= u) break;
     }
     std::vector<uint8_t> res(ids.size(), 0);
     uint64_t msk = 0, val = 0;
     for (size_t i = 0; i < ids.size(); ++i) {
          const size_t p = find_sure(ids[i]);
          msk |= uint64_t(1) << p;
          if ((k >> p) & 1) {
               val |= uint64_t(1) << p;
               res[i] = 1;
          }
     }
     const double norm = masked_norm(msk, val);
     normalize(norm, msk, val);
     return res;
}

double Engine::get_probability(const std::vector<uint8_t>& bits, const std::vector<int64_t>& ids)
{
     DeviceScope ds(device_);
     if (ids.size() != bits.size())
          throw std::runtime_error("GetProbability(): ids.size() != bit_string.size()");
     uint64_t msk, val;
     mask_of(ids, bits, msk, val);
     settle();
     return masked_norm(msk, val);
}

cplx Engine::get_amplitude(const std::vector<uint8_t>& bits, const std::vector<int64_t>& ids)
{
     DeviceScope ds(device_);
     if (bits.size() != locals_.size() || bits.size() != ids.size())
          throw std::runtime_error("GetAmplitude(): ids.size() != number of qubits");
     uint64_t num = 0, chk = 0;
     for (size_t i = 0; i < ids.size(); ++i) {
          const size_t p = find_sure(ids[i]);
          if (bits[i]) num |= uint64_t(1) << p;
          chk |= uint64_t(1) << p;
     }
     const uint64_t full = locals_.empty() ? 0 : ((uint64_t(1) << locals_.size()) - 1);
     if (chk != full)
          throw std::runtime_error(
              "GetAmplitude(): the second argument must be a permutation of all allocated "
              "qubits.");
     settle();
     const size_t ab = amp_bytes(prec_);
     char buf[16];
     check(hipMemcpyAsync(buf, static_cast<char*>(d_state_) + num * ab, ab, hipMemcpyDeviceToHost,
                          stream_),
           "hipMemcpyAsync(amplitude)");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     if (prec_ == Prec::C128) {
          double v[2];
          std::memcpy(v, buf, 16);
          return cplx(v[0], v[1]);
     }
     float v[2];
     std::memcpy(v, buf, 8);
     return cplx(v[0], v[1]);
}

void Engine::collapse(const std::vector<int64_t>& ids, const std::vector<uint8_t>& values)
{
     DeviceScope ds(device_);
     if (ids.size() != values.size())
          throw std::runtime_error("collapseWaveFunction(): ids.size() != values.size()");
     uint64_t msk, val;
     mask_of(ids, values, msk, val);
     settle();
     const double norm = masked_norm(msk, val);
     if (norm < 1.e-12)
          throw std::runtime_error("collapseWaveFunction(): Invalid collapse! Probability is ~0.");
     normalize(norm, msk, val);
}

double Engine::entropy()
{
     DeviceScope ds(device_);
     settle();
     begin_timed(kReduce);
     check(launch_entropy(d_state_, prec_, n_amps_, d_scratch_, d_scratch_ + kReduceScratchDoubles,
                          stream_),
           "entropy");
     end_timed();
     read_out(h_out_, 1);
     return -h_out_[0];
}

// ---------------------------------------------------------------------------
// State transfer
// ---------------------------------------------------------------------------
void* Engine::device_state()
{
     DeviceScope ds(device_);
     run();
     settle();
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     return d_state_;
}

void* Engine::pin_state()
{
     void* p = device_state();
     view_pins_.fetch_add(1);
     return p;
}

void Engine::unpin_state()
{
     int v = view_pins_.load();
     while (v > 0 && !view_pins_.compare_exchange_weak(v, v - 1)) {
     }
}

namespace {
constexpr char kStateMagic[4] = {'H', 'Q', 'S', 'V'};
constexpr uint32_t kStateVersion = 1;
constexpr size_t kStageBytes = size_t(64) << 20;

struct File {
     FILE* f = nullptr;
     File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode))
     {
          if (!f) throw std::runtime_error("state file '" + path + "': cannot open");
     }
     ~File()
     {
          if (f) std::fclose(f);
     }
     void read(void* p, size_t n)
     {
          if (std::fread(p, 1, n, f) != n) throw std::runtime_error("state file: truncated");
     }
     void write(const void* p, size_t n)
     {
          if (std::fwrite(p, 1, n, f) != n) throw std::runtime_error("state file: write failed");
     }
};

struct Pinned {
     void* p = nullptr;
     explicit Pinned(size_t n) { check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc(stage)"); }
     ~Pinned()
     {
          if (p) (void)hipHostFree(p);
     }
};
}  // namespace

void Engine::save_state(const std::string& path)
{
     DeviceScope ds(device_);
     run();
     settle();
     File f(path, "wb");
     const uint32_t hdr[3] = {kStateVersion, static_cast<uint32_t>(prec_), static_cast<uint32_t>(locals_.size())};
     f.write(kStateMagic, 4);
     f.write(hdr, sizeof(hdr));
     f.write(locals_.data(), locals_.size() * sizeof(int64_t));
     const size_t total = n_amps_ * amp_bytes(prec_);
     Pinned stage(std::min(total, kStageBytes));
     for (size_t off = 0; off < total; off += kStageBytes) {
          const size_t n = std::min(kStageBytes, total - off);
          check(hipMemcpyAsync(stage.p, static_cast<const char*>(d_state_) + off, n, hipMemcpyDeviceToHost, stream_),
                "hipMemcpyAsync(save)");
          check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
          f.write(stage.p, n);
     }
}

void Engine::load_state(const std::string& path)
{
     DeviceScope ds(device_);
     run();
     settle();
     File f(path, "rb");
     char magic[4];
     uint32_t hdr[3];
     f.read(magic, 4);
     f.read(hdr, sizeof(hdr));
     if (std::memcmp(magic, kStateMagic, 4) != 0 || hdr[0] != kStateVersion)
          throw std::runtime_error("state file '" + path + "': not a saved state");
     if (hdr[1] > 1) throw std::runtime_error("state file '" + path + "': unknown precision");
     const Prec fprec = static_cast<Prec>(hdr[1]);
     const uint32_t n = hdr[2];
     if (n != locals_.size())
          throw std::runtime_error("load_state(): the file holds " + std::to_string(n) + " qubits, the simulator " +
                                   std::to_string(locals_.size()));
     std::vector<int64_t> ids(n);
     f.read(ids.data(), n * sizeof(int64_t));
     {
          std::vector<int64_t> a = ids, b = locals_;
          std::sort(a.begin(), a.end());
          std::sort(b.begin(), b.end());
          if (a != b) throw std::runtime_error("load_state(): the file's qubit ids differ from the allocated ones");
     }
     const size_t fab = amp_bytes(fprec), ab = amp_bytes(prec_);
     {
          // the whole payload must be there before the state is touched
          const long here = std::ftell(f.f);
          if (here < 0 || std::fseek(f.f, 0, SEEK_END) != 0)
               throw std::runtime_error("state file '" + path + "': cannot seek");
          const long end = std::ftell(f.f);
          if (end < 0 || static_cast<uint64_t>(end - here) != n_amps_ * fab)
               throw std::runtime_error("state file '" + path + "': expected " + std::to_string(n_amps_ * fab) +
                                        " bytes of amplitudes");
          if (std::fseek(f.f, here, SEEK_SET) != 0) throw std::runtime_error("state file '" + path + "': cannot seek");
     }
     set_qubits_perm(ids);
     const uint64_t chunk = kStageBytes / std::max(fab, ab);
     Pinned stage(std::min<size_t>(n_amps_, chunk) * ab);
     std::vector<char> conv(fprec == prec_ ? 0 : std::min<size_t>(n_amps_, chunk) * fab);
     for (uint64_t a0 = 0; a0 < n_amps_; a0 += chunk) {
          const uint64_t na = std::min<uint64_t>(chunk, n_amps_ - a0);
          check(hipStreamSynchronize(stream_), "hipStreamSynchronize");  // stage free again
          if (fprec == prec_) {
               f.read(stage.p, na * ab);
          }
          else {
               f.read(conv.data(), na * fab);
               for (uint64_t i = 0; i < 2 * na; ++i) {
                    if (prec_ == Prec::C128)
                         static_cast<double*>(stage.p)[i] = reinterpret_cast<const float*>(conv.data())[i];
                    else
                         static_cast<float*>(stage.p)[i] =
                             static_cast<float>(reinterpret_cast<const double*>(conv.data())[i]);
               }
          }
          check(hipMemcpyAsync(static_cast<char*>(d_state_) + a0 * ab, stage.p, na * ab, hipMemcpyHostToDevice, stream_),
                "hipMemcpyAsync(load)");
     }
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::get_state_c128(std::complex<double>* out)
{
     DeviceScope ds(device_);
     settle();
     if (prec_ == Prec::C128) {
          check(hipMemcpyAsync(out, d_state_, n_amps_ * 16, hipMemcpyDeviceToHost, stream_),
                "hipMemcpyAsync(state)");
          check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
          return;
     }
     std::vector<std::complex<float>> tmp(n_amps_);
     check(hipMemcpyAsync(tmp.data(), d_state_, n_amps_ * 8, hipMemcpyDeviceToHost, stream_),
           "hipMemcpyAsync(state)");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
     for (uint64_t i = 0; i < n_amps_; ++i) out[i] = std::complex<double>(tmp[i]);
}

void Engine::set_state_c128(const std::complex<double>* in, uint64_t n)
{
     DeviceScope ds(device_);
     if (n != n_amps_)
          throw std::runtime_error("set_wavefunction(): expected " + std::to_string(n_amps_)
                                   + " amplitudes, got " + std::to_string(n));
     settle();
     if (prec_ == Prec::C128) {
          check(hipMemcpyAsync(d_state_, in, n * 16, hipMemcpyHostToDevice, stream_),
                "hipMemcpyAsync(state)");
          check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
          return;
     }
     std::vector<std::complex<float>> tmp(n);
     for (uint64_t i = 0; i < n; ++i) tmp[i] = std::complex<float>(in[i]);
     check(hipMemcpyAsync(d_state_, tmp.data(), n * 8, hipMemcpyHostToDevice, stream_),
           "hipMemcpyAsync(state)");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------
// Device-side checkpoint
// ---------------------------------------------------------------------------
void Engine::snapshot()
{
     DeviceScope ds(device_);
     settle();
     if (n_amps_ > cap_snap_) {
          if (d_snap_) {
               check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
               check(hipFree(d_snap_), "hipFree");
               d_snap_ = nullptr;
               cap_snap_ = 0;
          }
          check(hipMalloc(&d_snap_, n_amps_ * amp_bytes(prec_)), "hipMalloc(snapshot)");
          cap_snap_ = n_amps_;
     }
     begin_timed(kFillCopy);
     check(launch_copy(d_snap_, d_state_, prec_, n_amps_, stream_), "copy");
     end_timed();
     n_snap_ = n_amps_;
     snap_locals_ = locals_;
     has_snap_ = true;
}

void Engine::restore()
{
     DeviceScope ds(device_);
     if (!has_snap_) throw std::runtime_error("restore(): no snapshot taken");
     settle();
     collect_retired();
     if (n_snap_ > cap_state_) {
          release_buffer(d_state_);
          cap_state_ = 0;
          check(hipMalloc(&d_state_, n_snap_ * amp_bytes(prec_)), "hipMalloc(state)");
          cap_state_ = n_snap_;
     }
     begin_timed(kFillCopy);
     check(launch_copy(d_state_, d_snap_, prec_, n_snap_, stream_), "copy");
     end_timed();
     n_amps_ = n_snap_;
     locals_ = snap_locals_;
}

void Engine::compare_snapshot(double out[5])
{
     DeviceScope ds(device_);
     if (!has_snap_ || n_snap_ != n_amps_)
          throw std::runtime_error("compare_snapshot(): no snapshot of the current size");
     settle();
     begin_timed(kReduce);
     check(launch_compare(d_state_, d_snap_, prec_, n_amps_, d_scratch_,
                          d_scratch_ + kReduceScratchDoubles, stream_),
           "compare");
     end_timed();
     read_out(h_out_, 5);
     for (int i = 0; i < 5; ++i) out[i] = h_out_[i];
}

void Engine::init_random_state(uint64_t seed)
{
     DeviceScope ds(device_);
     settle();
     begin_timed(kFillCopy);
     check(launch_random_state(d_state_, prec_, n_amps_, seed, stream_), "random_state");
     end_timed();
     const double norm = masked_norm(0, 0);
     if (!(norm > 0)) throw std::runtime_error("init_random_state(): zero norm");
     const double f = 1.0 / std::sqrt(norm);
     begin_timed(kScale);
     check(launch_scale(d_state_, prec_, n_amps_, f, 0.0, stream_), "scale");
     end_timed();
}

int Engine::apply_tiled(const std::vector<Gate>& gates, int op_width)
{
     DeviceScope ds(device_);
     if (op_width < 1 || op_width > 5) throw std::runtime_error("apply_tiled(): op_width must be 1..5");
     run();
     settle();
     if (gates.empty()) return 0;
     const int n = static_cast<int>(n_qubits());
     // bit positions of every gate
     std::vector<PosGate> items;
     items.reserve(gates.size());
     for (const auto& g: gates) {
          if (g.targets.empty() || g.targets.size() + g.ctrls.size() > 5)
               throw std::runtime_error("apply_tiled(): a gate needs 1..5 qubits with its controls");
          if (g.m.dim != (size_t(1) << g.targets.size()))
               throw std::runtime_error("apply_tiled(): matrix size does not match targets");
          PosGate pg;
          for (auto id: g.targets) pg.t.push_back(static_cast<int64_t>(find_sure(id)));
          for (auto id: g.ctrls) pg.c.push_back(static_cast<int64_t>(find_sure(id)));
          std::vector<int64_t> all = pg.t;
          all.insert(all.end(), pg.c.begin(), pg.c.end());
          std::sort(all.begin(), all.end());
          if (std::adjacent_find(all.begin(), all.end()) != all.end())
               throw std::runtime_error("apply_tiled(): repeated qubit");
          pg.m = g.m;
          items.push_back(std::move(pg));
     }
     if (n >= kTileBits) return tile_items(items, op_width, 0);
     // small states: every op is one plain fused pass
     GateList gl;
     for (const auto& it: items) {
          gl.gate.push_back(it.t);
          gl.ctrl.push_back(it.c);
          gl.diag.push_back(false);
     }
     const std::vector<int> cl = plan_clusters(gl, op_width, kPlanLookahead);
     const int nc = *std::max_element(cl.begin(), cl.end()) + 1;
     std::vector<std::vector<int>> members(nc);
     for (size_t i = 0; i < cl.size(); ++i) members[cl[i]].push_back(static_cast<int>(i));
     for (const auto& mem: members) {
          Fusion f;
          for (int i: mem) {
               Mat m = items[i].m;
               std::vector<int64_t> ids = items[i].t;
               add_controls(m, ids, items[i].c);
               f.insert(std::move(m), false, std::move(ids), {});
          }
          Mat M;
          std::vector<int64_t> pos, ctrls;
          bool diag = false;
          f.perform(M, pos, ctrls, diag);
          std::vector<uint32_t> p32(pos.begin(), pos.end());
          dense_or_diag(M, p32, 0, false);
     }
     return nc;
}

int Engine::tile_items(std::vector<PosGate>& items, int op_width, size_t keep)
{
     const int n = n_qubits();
     if (n < kTileBits) throw std::runtime_error("tile_items(): the state is smaller than a tile");
     GateList gl;
     for (const auto& it: items) {
          gl.gate.push_back(it.t);
          gl.ctrl.push_back(it.c);
          gl.diag.push_back(false);
     }
     TilePassPlanner planner(gl, n, kTileHigh, op_width);
     std::vector<TilePlan> plans;
     {
          TilePlan tp;
          while (planner.remaining().size() > keep && planner.next(tp)) plans.push_back(std::move(tp));
     }
     if (plans.empty()) return 0;
     // host image: per pass its ops, then all matrices (fused in FP64)
     const size_t ab = amp_bytes(prec_);
     std::vector<TileOp> ops;
     std::vector<size_t> first;
     std::vector<int> kmax;
     std::vector<char> mats;
     uint64_t entries = 0;
     std::vector<char> used(items.size(), 0);
     for (const auto& tp: plans) {
          first.push_back(ops.size());
          int km = 1;
          int rank[64];
          for (int p = 0; p < 64; ++p) rank[p] = -1;
          for (int p = 0; p < 6; ++p) rank[p] = p;
          for (int j = 0; j < kTileHigh; ++j) rank[tp.hpos[j]] = 6 + j;
          for (const auto& op: tp.ops) {
               Fusion f;
               for (int i: op) {
                    used[i] = 1;
                    Mat m = items[i].m;
                    std::vector<int64_t> ids = items[i].t;
                    add_controls(m, ids, items[i].c);
                    f.insert(std::move(m), false, std::move(ids), {});
               }
               Mat M;
               std::vector<int64_t> pos, ctrls;
               bool diag = false;
               f.perform(M, pos, ctrls, diag);
               TileOp to{};
               to.k = static_cast<uint8_t>(pos.size());
               if (to.k < 1 || to.k > 5) throw std::runtime_error("apply_tiled(): op width out of range");
               for (size_t l = 0; l < pos.size(); ++l) {
                    if (rank[pos[l]] < 0) throw std::runtime_error("apply_tiled(): op outside its tile");
                    to.rpos[l] = static_cast<uint8_t>(rank[pos[l]]);
               }
               to.mat_off = entries;
               km = std::max(km, static_cast<int>(to.k));
               const size_t old = mats.size();
               mats.resize(old + M.a.size() * ab);
               put_tile_matrix(mats.data() + old, M.a.data(), to.k, prec_);
               entries += M.a.size();
               ops.push_back(to);
          }
          kmax.push_back(km);
     }
     const size_t ops_bytes = (ops.size() * sizeof(TileOp) + 255) / 256 * 256;
     const size_t need = ops_bytes + mats.size();
     // staging slot: waits only for the launches of kProgSlots groups ago
     ProgSlot& ps = prog_[prog_next_];
     prog_next_ = (prog_next_ + 1) % kProgSlots;
     if (ps.busy) check(hipEventSynchronize(ps.ev), "hipEventSynchronize");
     ps.busy = false;
     if (need > ps.cap) {
          if (ps.h) check(hipHostFree(ps.h), "hipHostFree");
          if (ps.d) check(hipFree(ps.d), "hipFree");
          ps.h = ps.d = nullptr;
          ps.cap = 0;
          const size_t cap = std::max(need, size_t(1) << 20);
          check(hipHostMalloc(reinterpret_cast<void**>(&ps.h), cap, hipHostMallocDefault), "hipHostMalloc(program)");
          check(hipMalloc(reinterpret_cast<void**>(&ps.d), cap), "hipMalloc(program)");
          ps.cap = cap;
     }
     if (!ps.ev) check(hipEventCreateWithFlags(&ps.ev, hipEventDisableTiming), "hipEventCreate");
     std::memcpy(ps.h, ops.data(), ops.size() * sizeof(TileOp));
     std::memcpy(ps.h + ops_bytes, mats.data(), mats.size());
     check(hipMemcpyAsync(ps.d, ps.h, need, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync(program)");
     const TileOp* d_ops = reinterpret_cast<const TileOp*>(ps.d);
     const void* d_mats = ps.d + ops_bytes;
     for (size_t p = 0; p < plans.size(); ++p) {
          TileHigh th{};
          for (int j = 0; j < kTileHigh; ++j) th.hpos[j] = static_cast<uint8_t>(plans[p].hpos[j]);
          const size_t nops = plans[p].ops.size();
          begin_timed(kTile);
          check(launch_tile_pass(d_state_, prec_, n, th, d_ops + first[p], static_cast<int>(nops), kmax[p],
                                 d_mats, stream_),
                "tile pass");
          end_timed();
     }
     check(hipEventRecord(ps.ev, stream_), "hipEventRecord");
     ps.busy = true;
     // the launched items leave the list; the rest keep program order
     std::vector<PosGate> rest;
     rest.reserve(items.size());
     for (size_t i = 0; i < items.size(); ++i)
          if (!used[i]) rest.push_back(std::move(items[i]));
     items.swap(rest);
     return static_cast<int>(plans.size());
}

void Engine::sample(uint32_t shots, uint64_t seed, uint64_t* out)
{
     DeviceScope ds(device_);
     if (shots == 0) return;
     if (shots > (1u << 20)) throw std::runtime_error("sample(): at most 2^20 shots");
     if (!out) throw std::runtime_error("sample(): null output");
     run();
     settle();
     int n = 0;
     while ((uint64_t(1) << n) < n_amps_) ++n;
     if (shots > samples_cap_) {
          check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
          if (d_samples_) check(hipFree(d_samples_), "hipFree");
          d_samples_ = nullptr;
          check(hipMalloc(reinterpret_cast<void**>(&d_samples_), shots * sizeof(uint64_t)), "hipMalloc(samples)");
          samples_cap_ = shots;
     }
     const uint32_t lb = sample_log_blocks(n);
     BatchItem it{};
     it.log_groups = static_cast<uint32_t>(n) - lb;
     SampleJob job{};
     job.seed = seed;
     job.log_blocks = lb;
     job.log_block_amps = static_cast<uint32_t>(n) - lb;
     job.shots = shots;
     // Descriptors go through the pinned output buffer into the output area
     // of the scratch (both hold 2 * kMaxOuts doubles).
     char* h = reinterpret_cast<char*>(h_out_);
     std::memcpy(h, &it, sizeof(it));
     std::memcpy(h + 256, &job, sizeof(job));
     char* d = reinterpret_cast<char*>(d_scratch_ + kReduceScratchDoubles);
     check(hipMemcpyAsync(d, h, 512, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync(descriptors)");
     begin_timed(kReduce);
     check(launch_block_sums_multi(d_state_, prec_, reinterpret_cast<const BatchItem*>(d), 1,
                                   uint64_t(1) << lb, d_blocks_, stream_),
           "block_sums_multi");
     check(launch_sample(d_state_, prec_, reinterpret_cast<const SampleJob*>(d + 256), 1, d_blocks_,
                         d_samples_, stream_),
           "sample");
     end_timed();
     check(hipMemcpyAsync(out, d_samples_, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_),
           "hipMemcpyAsync(samples)");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------
// Pauli sums
// ---------------------------------------------------------------------------
namespace {
inline cplx ipow(int e)
{
     switch (e & 3) {
          case 0: return cplx(1, 0);
          case 1: return cplx(0, 1);
          case 2: return cplx(-1, 0);
          default: return cplx(0, -1);
     }
}
}  // namespace

cplx Engine::expectation(const std::vector<uint64_t>& xm, const std::vector<uint64_t>& zm,
                         const std::vector<cplx>& coeff, const std::vector<int64_t>& ids)
{
     DeviceScope ds(device_);
     if (xm.size() != zm.size() || xm.size() != coeff.size())
          throw std::runtime_error("expectation(): term arrays differ in length");
     settle();
     std::vector<uint32_t> pos(ids.size());
     for (size_t i = 0; i < ids.size(); ++i) pos[i] = static_cast<uint32_t>(find_sure(ids[i]));
     auto to_bits = [&](uint64_t m) {
          uint64_t r = 0;
          for (size_t i = 0; i < 64; ++i)
               if ((m >> i) & 1) {
                    if (i >= ids.size())
                         throw std::runtime_error(
                             "qubit_operator acts on more qubits than contained in the qureg.");
                    r |= uint64_t(1) << pos[i];
               }
          return r;
     };
     cplx total(0.0, 0.0);
     for (size_t t0 = 0; t0 < xm.size(); t0 += kMaxOuts / 2) {
          const size_t t1 = std::min(xm.size(), t0 + kMaxOuts / 2);
          for (size_t t = t0; t < t1; ++t) {
               begin_timed(kPauli);
               check(launch_pauli_expectation(d_state_, prec_, n_amps_, to_bits(xm[t]),
                                              to_bits(zm[t]), d_scratch_,
                                              d_scratch_ + kReduceScratchDoubles + 2 * (t - t0),
                                              stream_),
                     "pauli_expectation");
               end_timed();
          }
          read_out(h_out_, 2 * (t1 - t0));
          for (size_t t = t0; t < t1; ++t) {
               const cplx s(h_out_[2 * (t - t0)], h_out_[2 * (t - t0) + 1]);
               total += coeff[t] * ipow(__builtin_popcountll(xm[t] & zm[t])) * s;
          }
     }
     return total;
}

void Engine::apply_pauli_sum(const std::vector<uint64_t>& xm, const std::vector<uint64_t>& zm,
                             const std::vector<cplx>& coeff, const std::vector<int64_t>& ids)
{
     DeviceScope ds(device_);
     if (xm.size() != zm.size() || xm.size() != coeff.size())
          throw std::runtime_error("apply_qubit_operator(): term arrays differ in length");
     std::vector<uint32_t> pos(ids.size());
     for (size_t i = 0; i < ids.size(); ++i) pos[i] = static_cast<uint32_t>(find_sure(ids[i]));
     auto to_bits = [&](uint64_t m) {
          uint64_t r = 0;
          for (size_t i = 0; i < 64; ++i)
               if ((m >> i) & 1) {
                    if (i >= ids.size())
                         throw std::runtime_error(
                             "qubit_operator acts on more qubits than contained in the qureg.");
                    r |= uint64_t(1) << pos[i];
               }
          return r;
     };
     std::vector<uint64_t> X(xm.size()), Z(xm.size());
     for (size_t t = 0; t < xm.size(); ++t) {
          X[t] = to_bits(xm[t]);
          Z[t] = to_bits(zm[t]);
     }
     settle();
     ensure_tmp(n_amps_);
     begin_timed(kFillCopy);
     check(launch_fill(d_tmp_, prec_, n_amps_, 0.0, 0.0, stream_), "fill");
     end_timed();
     for (size_t t = 0; t < xm.size(); ++t) {
          const cplx c = coeff[t] * ipow(__builtin_popcountll(xm[t] & zm[t]));
          begin_timed(kPauli);
          check(launch_pauli_accumulate(d_tmp_, d_state_, prec_, n_amps_, X[t], Z[t], c.real(),
                                        c.imag(), stream_),
                "pauli_accumulate");
          end_timed();
     }
     commit_tmp();
}

void Engine::time_evolution(const std::vector<uint64_t>& xm, const std::vector<uint64_t>& zm,
                            const std::vector<cplx>& coeff, double t, const std::vector<int64_t>& ids,
                            const std::vector<int64_t>& ctrls)
{
     DeviceScope ds(device_);
     if (xm.size() != zm.size() || xm.size() != coeff.size())
          throw std::runtime_error("emulate_time_evolution(): term arrays differ in length");
     if (!std::isfinite(t)) throw std::runtime_error("emulate_time_evolution(): time must be finite");
     run();
     settle();
     std::vector<uint32_t> pos(ids.size());
     for (size_t i = 0; i < ids.size(); ++i) pos[i] = static_cast<uint32_t>(find_sure(ids[i]));
     uint64_t cmask = 0, tmask = 0;
     for (auto p: pos) tmask |= uint64_t(1) << p;
     for (auto c: ctrls) cmask |= uint64_t(1) << find_sure(c);
     if (cmask & tmask) throw std::runtime_error("emulate_time_evolution(): a control is also a target");
     auto to_bits = [&](uint64_t m) {
          uint64_t r = 0;
          for (size_t i = 0; i < 64; ++i)
               if ((m >> i) & 1) {
                    if (i >= ids.size())
                         throw std::runtime_error("qubit_operator acts on more qubits than contained in the qureg.");
                    r |= uint64_t(1) << pos[i];
               }
          return r;
     };
     // identity terms -> phase; the others with i^{|x & z|} folded in (Y = i X Z)
     cplx c0(0, 0);
     double l1 = 0;
     std::vector<uint64_t> X, Z;
     std::vector<cplx> C;
     for (size_t k = 0; k < xm.size(); ++k) {
          if (xm[k] == 0 && zm[k] == 0) {
               c0 += coeff[k];
               continue;
          }
          X.push_back(to_bits(xm[k]));
          Z.push_back(to_bits(zm[k]));
          C.push_back(coeff[k] * ipow(__builtin_popcountll(xm[k] & zm[k])));
          l1 += std::abs(coeff[k]);
     }
     const size_t bytes = n_amps_ * amp_bytes(prec_);
     struct Buf {
          void* p = nullptr;
          explicit Buf(size_t n) { check(hipMalloc(&p, n), "hipMalloc(time evolution)"); }
          ~Buf()
          {
               if (p) (void)hipFree(p);
          }
     };
     // w: the controlled part of the state, evolved step by step
     Buf w(bytes), out(bytes), a(bytes), b(bytes);
     check(launch_copy(w.p, d_state_, prec_, n_amps_, stream_), "copy");
     if (cmask) {
          check(launch_normalize(w.p, prec_, n_amps_, cmask, cmask, 1.0, stream_), "select");
          // the state keeps only its uncontrolled part until w comes back
          check(launch_pauli_accumulate(d_state_, w.p, prec_, n_amps_, 0, 0, -1.0, 0.0, stream_), "subtract");
     }
     const uint64_t steps = X.empty() ? 0 : std::max<uint64_t>(1, static_cast<uint64_t>(std::ceil(std::abs(t) * l1)));
     const double tau = steps ? t / static_cast<double>(steps) : 0.0;
     constexpr int kMaxTerms = 64;
     void* pw = w.p;
     void* po = out.p;
     for (uint64_t s = 0; s < steps; ++s) {
          check(launch_copy(po, pw, prec_, n_amps_, stream_), "copy");
          void* pa = a.p;
          void* pb = b.p;
          check(launch_copy(pa, pw, prec_, n_amps_, stream_), "copy");
          for (int k = 1; k <= kMaxTerms; ++k) {
               // b = (-i tau / k) H a; out += b
               const cplx f = cplx(0, -tau / k);
               check(launch_fill(pb, prec_, n_amps_, 0.0, 0.0, stream_), "fill");
               for (size_t j = 0; j < X.size(); ++j) {
                    const cplx c = f * C[j];
                    check(launch_pauli_accumulate(pb, pa, prec_, n_amps_, X[j], Z[j], c.real(), c.imag(), stream_),
                          "pauli_accumulate");
               }
               check(launch_pauli_accumulate(po, pb, prec_, n_amps_, 0, 0, 1.0, 0.0, stream_), "accumulate");
               check(launch_masked_norm(pb, prec_, n_qubits(), 0, 0, d_scratch_, d_scratch_ + kReduceScratchDoubles,
                                        stream_),
                     "norm");
               read_out(h_out_, 1);
               std::swap(pa, pb);
               if (h_out_[0] < 1e-24) break;
          }
          std::swap(pw, po);
     }
     if (c0 != cplx(0, 0)) {
          const cplx ph = std::exp(cplx(0, -t) * c0);
          check(launch_scale(pw, prec_, n_amps_, ph.real(), ph.imag(), stream_), "phase");
     }
     if (cmask) check(launch_pauli_accumulate(d_state_, pw, prec_, n_amps_, 0, 0, 1.0, 0.0, stream_), "add");
     else check(launch_copy(d_state_, pw, prec_, n_amps_, stream_), "copy");
     check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

}  // namespace hq
