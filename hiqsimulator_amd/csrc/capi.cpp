// extern "C" boundary over hq::Engine (include/hiqsim_amd.h).
#include <complex>
#include <exception>
#include <string>
#include <vector>

#include "../../include/hiqsim_amd.h"
#include "engine.hpp"
#include "sched.hpp"

struct hiqsim_engine {
     hq::Engine eng;
     hiqsim_engine(uint64_t seed, int64_t ml, int64_t mc, hq::Prec p, int dev)
         : eng(seed, ml, mc, p, dev)
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

int hiqsim_version(void) { return 1; }

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

}  // extern "C"
