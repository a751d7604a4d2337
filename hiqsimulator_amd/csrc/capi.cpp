// extern "C" boundary over hq::Engine (include/hiqsim_amd.h).
#include <complex>
#include <exception>
#include <string>
#include <vector>

#include "../../include/hiqsim_amd.h"
#include "batch.hpp"
#include "engine.hpp"
#include "sched.hpp"
#include "noalloc.hpp"
#include "serve.hpp"

struct hiqsim_engine {
     hq::Engine eng;
     hiqsim_engine(uint64_t seed, int64_t ml, int64_t mc, hq::Prec p, int dev)
         : eng(seed, ml, mc, p, dev)
     {}
};

struct hiqsim_batch {
     hq::Batch b;
     hiqsim_batch(hq::Prec p, int dev) : b(p, dev) {}
};

struct hiqsim_server {
     hq::Server s;
     hiqsim_server(hq::Prec p, int dev, uint64_t sb, uint64_t mb, int pps, int ma)
         : s(p, dev, sb, mb, pps, ma)
     {}
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f)
{
     try {
          f();
          g_err.clear();
          return 0;
     } catch (const std::exception& e) {
          g_err = e.what();
     } catch (...) {
          g_err = "unknown error";
     }
     return 1;
}

void need(const void* p, const char* what)
{
     if (!p) throw std::runtime_error(std::string("null pointer: ") + what);
}

std::vector<int64_t> ivec(const int64_t* p, size_t n)
{
     if (n) need(p, "ids");
     return n ? std::vector<int64_t>(p, p + n) : std::vector<int64_t>();
}

std::vector<uint8_t> bvec(const uint8_t* p, size_t n)
{
     if (n) need(p, "bits");
     std::vector<uint8_t> v(n);
     for (size_t i = 0; i < n; ++i) v[i] = p[i] ? 1 : 0;
     return v;
}

hq::Mat mat(const double* m, size_t dim)
{
     need(m, "matrix");
     hq::Mat r(dim);
     for (size_t i = 0; i < dim * dim; ++i) r.a[i] = hq::cplx(m[2 * i], m[2 * i + 1]);
     return r;
}

void pauli_args(const uint64_t* x, const uint64_t* z, const double* c, size_t nt,
                std::vector<uint64_t>& X, std::vector<uint64_t>& Z, std::vector<hq::cplx>& C)
{
     if (nt) {
          need(x, "xmask");
          need(z, "zmask");
          need(c, "coeff");
     }
     X.assign(x, x + nt);
     Z.assign(z, z + nt);
     C.resize(nt);
     for (size_t t = 0; t < nt; ++t) C[t] = hq::cplx(c[2 * t], c[2 * t + 1]);
}

// Gate lists of the batch and request-loop entry points (CSR targets and
// controls, matrices back to back).
std::vector<hq::BatchGate> batch_gates(size_t ngates, const double* mats, const int64_t* tgt_ids,
                                       const size_t* tgt_off, const int64_t* ctrl_ids,
                                       const size_t* ctrl_off)
{
     std::vector<hq::BatchGate> gates(ngates);
     if (ngates) {
          need(mats, "mats");
          need(tgt_ids, "tgt_ids");
          need(tgt_off, "tgt_off");
          need(ctrl_off, "ctrl_off");
          if (ctrl_off[ngates] > ctrl_off[0]) need(ctrl_ids, "ctrl_ids");
     }
     size_t m = 0;
     for (size_t g = 0; g < ngates; ++g) {
          if (tgt_off[g + 1] < tgt_off[g] || ctrl_off[g + 1] < ctrl_off[g])
               throw std::runtime_error("gate offsets must be non-decreasing");
          const size_t nt = tgt_off[g + 1] - tgt_off[g];
          if (nt == 0 || nt > 5) throw std::runtime_error("a gate needs 1..5 targets");
          gates[g].targets.assign(tgt_ids + tgt_off[g], tgt_ids + tgt_off[g + 1]);
          gates[g].ctrls.assign(ctrl_ids + ctrl_off[g], ctrl_ids + ctrl_off[g + 1]);
          const size_t dd = (size_t(1) << nt) * (size_t(1) << nt);
          gates[g].m.resize(dd);
          for (size_t i = 0; i < dd; ++i) gates[g].m[i] = hq::cplx(mats[2 * (m + i)], mats[2 * (m + i) + 1]);
          m += dd;
     }
     return gates;
}

hq::GateList gate_list(const int64_t* gate_ids, const size_t* gate_off, const int64_t* ctrl_ids,
                       const size_t* ctrl_off, const uint8_t* diag, size_t ngates)
{
     need(gate_off, "gate_off");
     need(ctrl_off, "ctrl_off");
     if (ngates) need(diag, "diag");
     hq::GateList g;
     for (size_t i = 0; i < ngates; ++i) {
          if (gate_off[i + 1] < gate_off[i] || ctrl_off[i + 1] < ctrl_off[i])
               throw std::runtime_error("scheduler: offsets must be non-decreasing");
          g.gate.push_back(ivec(gate_ids + gate_off[i], gate_off[i + 1] - gate_off[i]));
          g.ctrl.push_back(ivec(ctrl_ids + ctrl_off[i], ctrl_off[i + 1] - ctrl_off[i]));
          g.diag.push_back(diag[i] != 0);
     }
     return g;
}

}  // namespace

extern "C" {

int hiqsim_version(void) { return 2; }

const char* hiqsim_last_error(void) { return g_err.c_str(); }

int hiqsim_device_count(int* out)
{
     return guard([&] {
          need(out, "out");
          int n = 0;
          if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
          *out = n;
     });
}

int hiqsim_create(uint64_t seed, int64_t max_local, int64_t max_cluster_size, int precision,
                  int device, hiqsim_t** out)
{
     return guard([&] {
          need(out, "out");
          if (precision != HIQSIM_C128 && precision != HIQSIM_C64)
               throw std::runtime_error("hiqsim_create(): unknown precision");
          *out = new hiqsim_engine(seed, max_local, max_cluster_size,
                                   precision == HIQSIM_C64 ? hq::Prec::C64 : hq::Prec::C128,
                                   device);
     });
}

int hiqsim_destroy(hiqsim_t* h)
{
     return guard([&] { delete h; });
}

int hiqsim_allocate_qubit(hiqsim_t* h, int64_t id)
{
     return guard([&] {
          need(h, "handle");
          h->eng.allocate_qubit(id);
     });
}

int hiqsim_allocate_qureg(hiqsim_t* h, const int64_t* ids, size_t n, double init_re,
                          double init_im)
{
     return guard([&] {
          need(h, "handle");
          h->eng.allocate_qureg(ivec(ids, n), hq::cplx(init_re, init_im));
     });
}

int hiqsim_deallocate_qubit(hiqsim_t* h, int64_t id)
{
     return guard([&] {
          need(h, "handle");
          h->eng.deallocate_qubit(id);
     });
}

int hiqsim_apply_controlled_gate(hiqsim_t* h, const double* m, size_t dim, const int64_t* ids,
                                 size_t nids, const int64_t* ctrls, size_t nctrls)
{
     return guard([&] {
          need(h, "handle");
          if (dim == 0 || dim > (size_t(1) << 20))
               throw std::runtime_error("apply_controlled_gate(): bad matrix dimension");
          h->eng.apply_gate(mat(m, dim), ivec(ids, nids), ivec(ctrls, nctrls));
     });
}

int hiqsim_run(hiqsim_t* h)
{
     return guard([&] {
          need(h, "handle");
          h->eng.run();
     });
}

int hiqsim_measure_qubits(hiqsim_t* h, const int64_t* ids, size_t n, uint8_t* out)
{
     return guard([&] {
          need(h, "handle");
          if (n) need(out, "out");
          auto r = h->eng.measure(ivec(ids, n));
          for (size_t i = 0; i < r.size(); ++i) out[i] = r[i];
     });
}

int hiqsim_get_probability(hiqsim_t* h, const uint8_t* bits, size_t nbits, const int64_t* ids,
                           size_t nids, double* out)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          *out = h->eng.get_probability(bvec(bits, nbits), ivec(ids, nids));
     });
}

int hiqsim_get_amplitude(hiqsim_t* h, const uint8_t* bits, size_t nbits, const int64_t* ids,
                         size_t nids, double* out)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          auto a = h->eng.get_amplitude(bvec(bits, nbits), ivec(ids, nids));
          out[0] = a.real();
          out[1] = a.imag();
     });
}

int hiqsim_collapse_wavefunction(hiqsim_t* h, const int64_t* ids, size_t nids,
                                 const uint8_t* values, size_t nvalues)
{
     return guard([&] {
          need(h, "handle");
          h->eng.collapse(ivec(ids, nids), bvec(values, nvalues));
     });
}

int hiqsim_entropy(hiqsim_t* h, double* out)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          *out = h->eng.entropy();
     });
}

int hiqsim_num_qubits(hiqsim_t* h, size_t* out)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          *out = h->eng.qubit_ids().size();
     });
}

int hiqsim_get_qubits_ids(hiqsim_t* h, int64_t* out, size_t cap, size_t* n)
{
     return guard([&] {
          need(h, "handle");
          need(n, "n");
          const auto& ids = h->eng.qubit_ids();
          *n = ids.size();
          if (cap < ids.size()) throw std::runtime_error("get_qubits_ids(): buffer too small");
          if (!ids.empty()) need(out, "out");
          for (size_t i = 0; i < ids.size(); ++i) out[i] = ids[i];
     });
}

int hiqsim_set_qubits_perm(hiqsim_t* h, const int64_t* ids, size_t n)
{
     return guard([&] {
          need(h, "handle");
          h->eng.set_qubits_perm(ivec(ids, n));
     });
}

int hiqsim_swap_qubits(hiqsim_t* h, const int64_t* pairs, size_t n)
{
     return guard([&] {
          need(h, "handle");
          h->eng.swap_qubits(ivec(pairs, n));
     });
}

int hiqsim_permute_qubits(hiqsim_t* h, const int64_t* ids, size_t n)
{
     return guard([&] {
          need(h, "handle");
          h->eng.permute_qubits(ivec(ids, n));
     });
}

int hiqsim_state_size(hiqsim_t* h, uint64_t* out)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          *out = h->eng.state_size();
     });
}

int hiqsim_device_state(hiqsim_t* h, void** ptr, uint64_t* n_amps, int* precision, int* device)
{
     return guard([&] {
          need(h, "handle");
          need(ptr, "ptr");
          need(n_amps, "n_amps");
          *ptr = h->eng.device_state();
          *n_amps = h->eng.state_size();
          if (precision) *precision = static_cast<int>(h->eng.precision());
          if (device) *device = h->eng.device();
     });
}

int hiqsim_pin_device_state(hiqsim_t* h, void** ptr, uint64_t* n_amps, int* precision, int* device)
{
     return guard([&] {
          need(h, "handle");
          need(ptr, "ptr");
          need(n_amps, "n_amps");
          *ptr = h->eng.pin_state();
          *n_amps = h->eng.state_size();
          if (precision) *precision = static_cast<int>(h->eng.precision());
          if (device) *device = h->eng.device();
     });
}

int hiqsim_unpin_device_state(hiqsim_t* h)
{
     return guard([&] {
          need(h, "handle");
          h->eng.unpin_state();
     });
}

int hiqsim_set_pass_mode(hiqsim_t* h, int mode)
{
     return guard([&] {
          need(h, "handle");
          h->eng.set_pass_mode(mode);
     });
}

int hiqsim_get_pass_mode(hiqsim_t* h, int* mode, size_t* queued)
{
     return guard([&] {
          need(h, "handle");
          need(mode, "mode");
          *mode = h->eng.pass_mode();
          if (queued) *queued = h->eng.queued_gates();
     });
}

int hiqsim_save_state(hiqsim_t* h, const char* path)
{
     return guard([&] {
          need(h, "handle");
          need(path, "path");
          h->eng.save_state(path);
     });
}

int hiqsim_load_state(hiqsim_t* h, const char* path)
{
     return guard([&] {
          need(h, "handle");
          need(path, "path");
          h->eng.load_state(path);
     });
}

int hiqsim_get_state(hiqsim_t* h, double* out, uint64_t n_amps)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          if (n_amps != h->eng.state_size())
               throw std::runtime_error("get_state(): n_amps does not match the state size");
          h->eng.get_state_c128(reinterpret_cast<std::complex<double>*>(out));
     });
}

int hiqsim_set_state(hiqsim_t* h, const double* in, uint64_t n_amps)
{
     return guard([&] {
          need(h, "handle");
          need(in, "in");
          h->eng.set_state_c128(reinterpret_cast<const std::complex<double>*>(in), n_amps);
     });
}

int hiqsim_snapshot_state(hiqsim_t* h)
{
     return guard([&] {
          need(h, "handle");
          h->eng.snapshot();
     });
}

int hiqsim_restore_state(hiqsim_t* h)
{
     return guard([&] {
          need(h, "handle");
          h->eng.restore();
     });
}

int hiqsim_compare_snapshot(hiqsim_t* h, double* out)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          h->eng.compare_snapshot(out);
     });
}

int hiqsim_init_random_state(hiqsim_t* h, uint64_t seed)
{
     return guard([&] {
          need(h, "handle");
          h->eng.init_random_state(seed);
     });
}

int hiqsim_expectation_pauli(hiqsim_t* h, const uint64_t* xmask, const uint64_t* zmask,
                             const double* coeff, size_t nterms, const int64_t* ids, size_t nids,
                             double* out)
{
     return guard([&] {
          need(h, "handle");
          need(out, "out");
          std::vector<uint64_t> X, Z;
          std::vector<hq::cplx> C;
          pauli_args(xmask, zmask, coeff, nterms, X, Z, C);
          auto r = h->eng.expectation(X, Z, C, ivec(ids, nids));
          out[0] = r.real();
          out[1] = r.imag();
     });
}

int hiqsim_apply_pauli_sum(hiqsim_t* h, const uint64_t* xmask, const uint64_t* zmask,
                           const double* coeff, size_t nterms, const int64_t* ids, size_t nids)
{
     return guard([&] {
          need(h, "handle");
          std::vector<uint64_t> X, Z;
          std::vector<hq::cplx> C;
          pauli_args(xmask, zmask, coeff, nterms, X, Z, C);
          h->eng.apply_pauli_sum(X, Z, C, ivec(ids, nids));
     });
}

int hiqsim_emulate_time_evolution(hiqsim_t* h, const uint64_t* xmask, const uint64_t* zmask,
                                  const double* coeff, size_t nterms, double time, const int64_t* ids,
                                  size_t nids, const int64_t* ctrls, size_t nctrls)
{
     return guard([&] {
          need(h, "handle");
          std::vector<uint64_t> X, Z;
          std::vector<hq::cplx> C;
          pauli_args(xmask, zmask, coeff, nterms, X, Z, C);
          h->eng.time_evolution(X, Z, C, time, ivec(ids, nids), ivec(ctrls, nctrls));
     });
}

int hiqsim_apply_matrix(hiqsim_t* h, const double* m, int k, const uint32_t* pos,
                        uint64_t ctrlmask, int diag)
{
     return guard([&] {
          need(h, "handle");
          if (k < 1 || k > 5) throw std::runtime_error("apply_matrix(): k must be 1..5");
          need(pos, "pos");
          h->eng.apply_matrix(mat(m, size_t(1) << k), std::vector<uint32_t>(pos, pos + k),
                              ctrlmask, diag != 0);
     });
}

int hiqsim_synchronize(hiqsim_t* h)
{
     return guard([&] {
          need(h, "handle");
          h->eng.synchronize();
     });
}

int hiqsim_set_timing(hiqsim_t* h, int on)
{
     return guard([&] {
          need(h, "handle");
          h->eng.set_timing(on != 0);
     });
}

int hiqsim_kernel_time(hiqsim_t* h, int cls, uint64_t* count, double* total_ms)
{
     return guard([&] {
          need(h, "handle");
          need(count, "count");
          need(total_ms, "total_ms");
          h->eng.kernel_time(cls, *count, *total_ms);
     });
}

int hiqsim_schedule_cluster(const int64_t* gate_ids, const size_t* gate_off,
                            const int64_t* ctrl_ids, const size_t* ctrl_off, const uint8_t* diag,
                            size_t ngates, const int64_t* locals, size_t nlocals,
                            const int64_t* globals, size_t nglobals, int cluster_size,
                            int32_t* out, size_t* nout)
{
     return guard([&] {
          need(nout, "nout");
          auto g = gate_list(gate_ids, gate_off, ctrl_ids, ctrl_off, diag, ngates);
          auto r = hq::schedule_cluster(g, ivec(locals, nlocals), ivec(globals, nglobals),
                                        cluster_size);
          if (!r.empty()) need(out, "out");
          for (size_t i = 0; i < r.size(); ++i) out[i] = r[i];
          *nout = r.size();
     });
}

int hiqsim_schedule_swap(const int64_t* gate_ids, const size_t* gate_off, const int64_t* ctrl_ids,
                         const size_t* ctrl_off, const uint8_t* diag, size_t ngates,
                         int64_t num_splits, int num_locals, int fuse, int64_t* out, size_t* nout)
{
     return guard([&] {
          need(nout, "nout");
          auto g = gate_list(gate_ids, gate_off, ctrl_ids, ctrl_off, diag, ngates);
          auto r = hq::schedule_swap(g, num_splits, num_locals, fuse != 0);
          if (!r.empty()) need(out, "out");
          for (size_t i = 0; i < r.size(); ++i) out[i] = r[i];
          *nout = r.size();
     });
}

int hiqsim_plan_clusters(const int64_t* gate_ids, const size_t* gate_off, const int64_t* ctrl_ids,
                         const size_t* ctrl_off, size_t ngates, int cluster_size, int mode,
                         int32_t* cluster_of)
{
     return guard([&] {
          if (ngates) need(cluster_of, "cluster_of");
          std::vector<uint8_t> diag(ngates, 0);
          auto g = gate_list(gate_ids, gate_off, ctrl_ids, ctrl_off, diag.data(), ngates);
          auto r = hq::plan_clusters(g, cluster_size, mode);
          for (size_t i = 0; i < r.size(); ++i) cluster_of[i] = r[i];
     });
}

int hiqsim_apply_tiled(hiqsim_t* h, size_t ngates, const double* mats, const int64_t* tgt_ids,
                       const size_t* tgt_off, const int64_t* ctrl_ids, const size_t* ctrl_off,
                       int op_width, int* npasses)
{
     return guard([&] {
          need(h, "handle");
          auto bg = batch_gates(ngates, mats, tgt_ids, tgt_off, ctrl_ids, ctrl_off);
          std::vector<hq::Engine::Gate> gates(bg.size());
          for (size_t i = 0; i < bg.size(); ++i) {
               const size_t dim = size_t(1) << bg[i].targets.size();
               gates[i].m = hq::Mat(dim);
               gates[i].m.a = std::move(bg[i].m);
               gates[i].targets = std::move(bg[i].targets);
               gates[i].ctrls = std::move(bg[i].ctrls);
          }
          const int np = h->eng.apply_tiled(gates, op_width);
          if (npasses) *npasses = np;
     });
}

int hiqsim_plan_tile_passes(const int64_t* gate_pos, const size_t* gate_off, const int64_t* ctrl_pos,
                            const size_t* ctrl_off, size_t ngates, int n_bits, int op_width,
                            int32_t* pass_of, int32_t* op_of, int32_t* hpos, size_t* npasses)
{
     return guard([&] {
          std::vector<uint8_t> diag(ngates, 0);
          auto g = gate_list(gate_pos, gate_off, ctrl_pos, ctrl_off, diag.data(), ngates);
          const auto plans = hq::plan_tile_passes(g, n_bits, hq::kTileHigh, op_width);
          if (ngates) {
               need(pass_of, "pass_of");
               need(op_of, "op_of");
               need(hpos, "hpos");
          }
          for (size_t p = 0; p < plans.size(); ++p) {
               for (int j = 0; j < hq::kTileHigh; ++j) hpos[hq::kTileHigh * p + j] = plans[p].hpos[j];
               for (size_t o = 0; o < plans[p].ops.size(); ++o)
                    for (int gi: plans[p].ops[o]) {
                         pass_of[gi] = static_cast<int32_t>(p);
                         op_of[gi] = static_cast<int32_t>(o);
                    }
          }
          if (npasses) *npasses = plans.size();
     });
}

int hiqsim_set_tuning(const char* name, int64_t value)
{
     return guard([&] {
          need(name, "name");
          const std::string n(name);
          if (n == "tile_high_target") {
               if (value < 0 || value > 64) throw std::runtime_error("tile_high_target must be 0..64");
               hq::set_tile_high_target(static_cast<int>(value));
          }
          else if (n == "block_swizzle") {
               if (value < 0 || value > 16) throw std::runtime_error("block_swizzle must be 0..16");
               hq::set_block_swizzle(static_cast<int>(value));
          }
          else if (n == "block_swizzle_wide") {
               if (value < 0 || value > 16) throw std::runtime_error("block_swizzle_wide must be 0..16");
               hq::set_block_swizzle_wide(static_cast<int>(value));
          }
          else if (n == "tile_mfma") {
               if (value < 0 || value > 1) throw std::runtime_error("tile_mfma must be 0 or 1");
               hq::set_tile_mfma(static_cast<int>(value));
          }
          else if (n == "dense5_tile") {
               if (value < 0 || value > 1) throw std::runtime_error("dense5_tile must be 0 or 1");
               hq::set_dense5_tile(static_cast<int>(value));
          }
          else if (n == "tile_high_target_wide") {
               if (value < 0 || value > 64) throw std::runtime_error("tile_high_target_wide must be 0..64");
               hq::set_tile_high_target_wide(static_cast<int>(value));
          }
          else if (n == "direct_block") {
               if (value != 256 && value != 512 && value != 1024)
                    throw std::runtime_error("direct_block must be 256, 512 or 1024");
               hq::set_direct_block(static_cast<int>(value));
          }
          else if (n == "permute_swizzle") {
               if (value < 0 || value > 16) throw std::runtime_error("permute_swizzle must be 0..16");
               hq::set_permute_swizzle(static_cast<int>(value));
          }
          else if (n == "tile_threads") {
               if (value != 0 && value != 256 && value != 512)
                    throw std::runtime_error("tile_threads must be 0 (per precision), 256 or 512");
               hq::set_tile_threads(static_cast<int>(value));
          }
          else if (n == "permute_block") {
               if (value != 256 && value != 512) throw std::runtime_error("permute_block must be 256 or 512");
               hq::set_permute_block(static_cast<int>(value));
          }
          else
               throw std::runtime_error("unknown tuning parameter '" + n + "'");
     });
}

int hiqsim_get_tuning(const char* name, int64_t* value)
{
     return guard([&] {
          need(name, "name");
          need(value, "value");
          const std::string n(name);
          if (n == "tile_high_target") *value = hq::get_tile_high_target();
          else if (n == "block_swizzle") *value = hq::get_block_swizzle();
          else if (n == "block_swizzle_wide") *value = hq::get_block_swizzle_wide();
          else if (n == "tile_mfma") *value = hq::get_tile_mfma();
          else if (n == "dense5_tile") *value = hq::get_dense5_tile();
          else if (n == "direct_block") *value = hq::get_direct_block();
          else if (n == "tile_high_target_wide") *value = hq::get_tile_high_target_wide();
          else if (n == "permute_swizzle") *value = hq::get_permute_swizzle();
          else if (n == "permute_block") *value = hq::get_permute_block();
          else if (n == "tile_threads") *value = hq::get_tile_threads();
          else throw std::runtime_error("unknown tuning parameter '" + n + "'");
     });
}

int hiqsim_batch_create(int precision, int device, hiqsim_batch_t** out)
{
     return guard([&] {
          need(out, "out");
          if (precision != HIQSIM_C128 && precision != HIQSIM_C64)
               throw std::runtime_error("precision must be HIQSIM_C128 or HIQSIM_C64");
          *out = new hiqsim_batch(static_cast<hq::Prec>(precision), device);
     });
}

int hiqsim_batch_destroy(hiqsim_batch_t* b)
{
     return guard([&] { delete b; });
}

int hiqsim_batch_add_circuit(hiqsim_batch_t* b, int n_qubits, size_t ngates, const double* mats,
                             const int64_t* tgt_ids, const size_t* tgt_off,
                             const int64_t* ctrl_ids, const size_t* ctrl_off, int64_t* index)
{
     return guard([&] {
          need(b, "batch");
          need(index, "index");
          *index = b->b.add_circuit(n_qubits, batch_gates(ngates, mats, tgt_ids, tgt_off, ctrl_ids,
                                                          ctrl_off));
     });
}

int hiqsim_batch_add_copies(hiqsim_batch_t* b, int64_t index, int64_t count, int64_t* first)
{
     return guard([&] {
          need(b, "batch");
          const int64_t f = b->b.add_copies(index, count);
          if (first) *first = f;
     });
}

int hiqsim_batch_set_plan(hiqsim_batch_t* b, const int32_t* k_by_n, size_t len, int wg_max_qubits,
                          int tile_min_qubits, int prog_min_qubits)
{
     return guard([&] {
          need(b, "batch");
          if (len) need(k_by_n, "k_by_n");
          b->b.set_plan(std::vector<int>(k_by_n, k_by_n + len), wg_max_qubits, tile_min_qubits,
                        prog_min_qubits);
     });
}

int hiqsim_batch_prepare(hiqsim_batch_t* b)
{
     return guard([&] {
          need(b, "batch");
          b->b.prepare();
     });
}

int hiqsim_batch_build(hiqsim_batch_t* b)
{
     return guard([&] {
          need(b, "batch");
          b->b.build();
     });
}

int hiqsim_batch_run(hiqsim_batch_t* b, double* device_ms)
{
     return guard([&] {
          need(b, "batch");
          const double ms = b->b.run();
          if (device_ms) *device_ms = ms;
     });
}

int hiqsim_batch_num_circuits(hiqsim_batch_t* b, int64_t* n)
{
     return guard([&] {
          need(b, "batch");
          need(n, "n");
          *n = b->b.num_circuits();
     });
}

int hiqsim_batch_num_qubits(hiqsim_batch_t* b, int64_t index, int* n_qubits)
{
     return guard([&] {
          need(b, "batch");
          need(n_qubits, "n_qubits");
          *n_qubits = b->b.num_qubits(index);
     });
}

int hiqsim_batch_get_state(hiqsim_batch_t* b, int64_t index, double* out)
{
     return guard([&] {
          need(b, "batch");
          need(out, "out");
          b->b.get_state(index, reinterpret_cast<std::complex<double>*>(out));
     });
}

int hiqsim_batch_stats(hiqsim_batch_t* b, uint64_t* out, size_t len)
{
     return guard([&] {
          need(b, "batch");
          if (len) need(out, "out");
          const hq::BatchStats& s = b->b.stats();
          const uint64_t v[14] = {s.circuits,       s.passes,      s.wg_circuits,
                                  s.lds_circuits,   s.layer_launches, s.arena_bytes,
                                  s.matrix_bytes,   s.amp_passes,  static_cast<uint64_t>(s.wg_grid),
                                  static_cast<uint64_t>(s.kmax), s.padded_launches, s.tile_passes,
                                  s.tile_ops,       s.tile_mfma_flops};
          for (size_t i = 0; i < len && i < 14; ++i) out[i] = v[i];
     });
}

int hiqsim_sample(hiqsim_t* h, uint32_t shots, uint64_t seed, uint64_t* out)
{
     return guard([&] {
          need(h, "handle");
          if (shots) need(out, "out");
          h->eng.sample(shots, seed, out);
     });
}

int hiqsim_sample_noalloc(const void* d_state, int n_qubits, int precision, uint32_t shots,
                          uint64_t seed, uint64_t* d_out, void* stream)
{
     return guard([&] {
          if (precision != HIQSIM_C128 && precision != HIQSIM_C64)
               throw std::runtime_error("precision must be HIQSIM_C128 or HIQSIM_C64");
          hq::sample_noalloc(d_state, n_qubits, static_cast<hq::Prec>(precision), shots, seed, d_out,
                             static_cast<hipStream_t>(stream));
     });
}

int hiqsim_server_create(int precision, int device, uint64_t state_bytes, uint64_t matrix_bytes,
                         int passes_per_step, int max_active, hiqsim_server_t** out)
{
     return guard([&] {
          need(out, "out");
          if (precision != HIQSIM_C128 && precision != HIQSIM_C64)
               throw std::runtime_error("precision must be HIQSIM_C128 or HIQSIM_C64");
          *out = new hiqsim_server(static_cast<hq::Prec>(precision), device, state_bytes, matrix_bytes,
                                   passes_per_step, max_active);
     });
}

int hiqsim_server_destroy(hiqsim_server_t* s)
{
     return guard([&] { delete s; });
}

int hiqsim_server_set_widths(hiqsim_server_t* s, const int32_t* k_by_n, size_t len)
{
     return guard([&] {
          need(s, "server");
          if (len) need(k_by_n, "k_by_n");
          s->s.set_widths(std::vector<int>(k_by_n, k_by_n + len));
     });
}

int hiqsim_server_set_prog_min_qubits(hiqsim_server_t* s, int q)
{
     return guard([&] {
          need(s, "server");
          s->s.set_prog_min_qubits(q);
     });
}

int hiqsim_server_set_heavy_qubits(hiqsim_server_t* s, int q)
{
     return guard([&] {
          need(s, "server");
          s->s.set_heavy_qubits(q);
     });
}

int hiqsim_server_submit(hiqsim_server_t* s, int n_qubits, size_t ngates, const double* mats,
                         const int64_t* tgt_ids, const size_t* tgt_off, const int64_t* ctrl_ids,
                         const size_t* ctrl_off, int kind, uint32_t shots, uint64_t seed,
                         const uint64_t* xmask, const uint64_t* zmask, size_t nterms, int64_t* id)
{
     return guard([&] {
          need(s, "server");
          need(id, "id");
          if (kind < 0 || kind > 2)
               throw std::runtime_error("kind must be 0 (samples), 1 (expectation) or 2 (state)");
          std::vector<uint64_t> X, Z;
          if (kind == 1 && nterms) {
               need(xmask, "xmask");
               need(zmask, "zmask");
               X.assign(xmask, xmask + nterms);
               Z.assign(zmask, zmask + nterms);
          }
          *id = s->s.submit(n_qubits, batch_gates(ngates, mats, tgt_ids, tgt_off, ctrl_ids, ctrl_off),
                            static_cast<hq::ServeKind>(kind), shots, seed, std::move(X), std::move(Z));
     });
}

int hiqsim_server_step(hiqsim_server_t* s, int block, int64_t* completed, size_t cap,
                       size_t* ncompleted)
{
     return guard([&] {
          need(s, "server");
          const auto done = s->s.step(block != 0);
          if (cap) need(completed, "completed");
          for (size_t i = 0; i < done.size() && i < cap; ++i) completed[i] = done[i];
          if (ncompleted) *ncompleted = done.size();
     });
}

int hiqsim_server_idle(hiqsim_server_t* s, int* idle)
{
     return guard([&] {
          need(s, "server");
          need(idle, "idle");
          *idle = s->s.idle() ? 1 : 0;
     });
}

int hiqsim_server_done(hiqsim_server_t* s, int64_t id, int* done)
{
     return guard([&] {
          need(s, "server");
          need(done, "done");
          *done = s->s.done(id) ? 1 : 0;
     });
}

int hiqsim_server_samples(hiqsim_server_t* s, int64_t id, uint64_t* out, size_t cap, size_t* n)
{
     return guard([&] {
          need(s, "server");
          const auto& v = s->s.samples(id);
          if (cap) need(out, "out");
          for (size_t i = 0; i < v.size() && i < cap; ++i) out[i] = v[i];
          if (n) *n = v.size();
     });
}

int hiqsim_server_expectations(hiqsim_server_t* s, int64_t id, double* out, size_t cap_terms,
                               size_t* n)
{
     return guard([&] {
          need(s, "server");
          const auto& v = s->s.expectations(id);
          if (cap_terms) need(out, "out");
          for (size_t i = 0; i < v.size() && i < cap_terms; ++i) {
               out[2 * i] = v[i].real();
               out[2 * i + 1] = v[i].imag();
          }
          if (n) *n = v.size();
     });
}

int hiqsim_server_state(hiqsim_server_t* s, int64_t id, double* out, uint64_t cap_amps, uint64_t* n_amps)
{
     return guard([&] {
          need(s, "server");
          const auto& v = s->s.state(id);
          if (cap_amps) need(out, "out");
          for (uint64_t i = 0; i < v.size() && i < cap_amps; ++i) {
               out[2 * i] = v[i].real();
               out[2 * i + 1] = v[i].imag();
          }
          if (n_amps) *n_amps = v.size();
     });
}

int hiqsim_server_times(hiqsim_server_t* s, int64_t id, double* out)
{
     return guard([&] {
          need(s, "server");
          need(out, "out");
          s->s.times(id, out, out + 1, out + 2);
     });
}

int hiqsim_server_forget(hiqsim_server_t* s, int64_t id)
{
     return guard([&] {
          need(s, "server");
          s->s.forget(id);
     });
}

int hiqsim_server_stats(hiqsim_server_t* s, uint64_t* out, size_t len)
{
     return guard([&] {
          need(s, "server");
          if (len) need(out, "out");
          const hq::ServeStats& t = s->s.stats();
          const uint64_t v[6] = {t.submitted, t.completed, t.steps, t.launches, t.peak_active,
                                 t.peak_state_bytes};
          for (size_t i = 0; i < len && i < 6; ++i) out[i] = v[i];
     });
}

}  // extern "C"
