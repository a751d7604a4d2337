// pybind11 module `_cppsim_amd`: the Python face of the engine, method for
// method the reference's `_cppsim_mpi.SimulatorMPI` (_cppsim_mpi.cpp:61-83).
// It calls only the C ABI (include/hiqsim_amd.h); errors become RuntimeError
// like the reference's std::runtime_error.  Device work runs with the GIL
// released.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <complex>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "hiqsim_amd.h"

namespace py = pybind11;

namespace {

// DLPack exchange structs (dlpack.h, the published DLPack ABI that
// torch.from_dlpack and numpy read), declared here from the specification.
struct DLDevice {
     int32_t device_type;
     int32_t device_id;
};
struct DLDataType {
     uint8_t code;
     uint8_t bits;
     uint16_t lanes;
};
struct DLTensor {
     void* data;
     DLDevice device;
     int32_t ndim;
     DLDataType dtype;
     int64_t* shape;
     int64_t* strides;
     uint64_t byte_offset;
};
struct DLManagedTensor {
     DLTensor dl_tensor;
     void* manager_ctx;
     void (*deleter)(DLManagedTensor*);
};
constexpr int32_t kDLROCM = 10;
constexpr uint8_t kDLComplex = 5;

// A DLPack view of a simulator's device state; it keeps the Python
// simulator object alive until the consumer releases the view.
struct StateView {
     DLManagedTensor t{};
     int64_t shape[1] = {0};
     PyObject* owner = nullptr;  // the SimulatorMPI object: keeps the handle alive
     hiqsim_t* handle = nullptr;
};

void state_view_deleter(DLManagedTensor* t)
{
     auto* v = static_cast<StateView*>(t->manager_ctx);
     if (v->handle) hiqsim_unpin_device_state(v->handle);  // only a counter
     {
          py::gil_scoped_acquire g;
          Py_XDECREF(v->owner);
     }
     delete v;
}

using cvec = std::vector<std::complex<double>>;

void ok(int rc)
{
     if (rc != 0) throw std::runtime_error(hiqsim_last_error());
}

std::vector<uint8_t> to_bits(const std::vector<bool>& b)
{
     std::vector<uint8_t> r(b.size());
     for (size_t i = 0; i < b.size(); ++i) r[i] = b[i] ? 1 : 0;
     return r;
}

// ProjectQ QubitOperator terms: [([(index, 'X'|'Y'|'Z'), ...], coeff), ...].
void parse_terms(const py::list& terms, std::vector<uint64_t>& X, std::vector<uint64_t>& Z,
                 std::vector<double>& C)
{
     for (auto item: terms) {
          auto tc = item.cast<py::sequence>();
          if (py::len(tc) != 2) throw std::runtime_error("terms must be (term, coeff) pairs");
          uint64_t x = 0, z = 0;
          for (auto op: tc[0].cast<py::sequence>()) {
               auto pr = op.cast<py::sequence>();
               const int64_t idx = pr[0].cast<int64_t>();
               const std::string p = pr[1].cast<std::string>();
               if (idx < 0 || idx >= 64) throw std::runtime_error("Pauli index out of range");
               const uint64_t bit = uint64_t(1) << idx;
               if ((x | z) & bit) throw std::runtime_error("repeated qubit in a Pauli term");
               if (p == "X") x |= bit;
               else if (p == "Y") {
                    x |= bit;
                    z |= bit;
               }
               else if (p == "Z") z |= bit;
               else throw std::runtime_error("unknown Pauli operator '" + p + "'");
          }
          const std::complex<double> c = tc[1].cast<std::complex<double>>();
          X.push_back(x);
          Z.push_back(z);
          C.push_back(c.real());
          C.push_back(c.imag());
     }
}

using CArr = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>;

// [(matrix, targets, controls), ...] in the CSR form of the batch entry points.
struct GateCsr {
     std::vector<double> mats;
     std::vector<int64_t> tid, cid;
     std::vector<size_t> toff{0}, coff{0};
     size_t size() const { return toff.size() - 1; }
};

GateCsr gate_csr(const py::list& gates)
{
     GateCsr g;
     for (auto item: gates) {
          auto t = item.cast<py::sequence>();
          if (py::len(t) != 3) throw std::runtime_error("gates must be (matrix, targets, controls)");
          CArr m = t[0].cast<CArr>();
          for (auto q: t[1].cast<py::sequence>()) g.tid.push_back(q.cast<int64_t>());
          for (auto q: t[2].cast<py::sequence>()) g.cid.push_back(q.cast<int64_t>());
          const size_t nt = g.tid.size() - g.toff.back();
          if (nt == 0 || nt > 5) throw std::runtime_error("a gate needs 1..5 targets");
          const size_t dd = (size_t(1) << nt) * (size_t(1) << nt);
          if (static_cast<size_t>(m.size()) != dd)
               throw std::runtime_error("gate matrix size does not match its targets");
          const std::complex<double>* d = m.data();
          for (size_t i = 0; i < dd; ++i) {
               g.mats.push_back(d[i].real());
               g.mats.push_back(d[i].imag());
          }
          g.toff.push_back(g.tid.size());
          g.coff.push_back(g.cid.size());
     }
     return g;
}

class PySim {
public:
     PySim(uint64_t seed, int64_t max_local, int64_t max_cluster, const std::string& precision,
           int device)
     {
          int p;
          if (precision == "c128" || precision == "complex128") p = HIQSIM_C128;
          else if (precision == "c64" || precision == "complex64") p = HIQSIM_C64;
          else throw std::runtime_error("precision must be 'c128' or 'c64'");
          ok(hiqsim_create(seed, max_local, max_cluster, p, device, &h_));
     }
     ~PySim()
     {
          if (h_) hiqsim_destroy(h_);
     }

     std::vector<int64_t> qubit_ids() const
     {
          size_t n = 0;
          ok(hiqsim_num_qubits(h_, &n));
          std::vector<int64_t> ids(n);
          ok(hiqsim_get_qubits_ids(h_, ids.data(), n, &n));
          return ids;
     }
     std::vector<int64_t> global_ids() const { return {}; }
     void set_qubits_perm(const std::vector<int64_t>& ids)
     {
          ok(hiqsim_set_qubits_perm(h_, ids.data(), ids.size()));
     }
     void swap_qubits(const std::vector<int64_t>& pairs)
     {
          ok(hiqsim_swap_qubits(h_, pairs.data(), pairs.size()));
     }
     void permute_qubits(const std::vector<int64_t>& ids)
     {
          py::gil_scoped_release r;
          ok(hiqsim_permute_qubits(h_, ids.data(), ids.size()));
     }
     void allocate_qureg(const std::vector<int64_t>& ids, std::complex<double> init)
     {
          py::gil_scoped_release r;
          ok(hiqsim_allocate_qureg(h_, ids.data(), ids.size(), init.real(), init.imag()));
     }
     void allocate_qubit(int64_t id)
     {
          py::gil_scoped_release r;
          ok(hiqsim_allocate_qubit(h_, id));
     }
     void deallocate_qubit(int64_t id)
     {
          py::gil_scoped_release r;
          ok(hiqsim_deallocate_qubit(h_, id));
     }
     std::vector<bool> measure_qubits(const std::vector<int64_t>& ids)
     {
          std::vector<uint8_t> out(ids.size());
          {
               py::gil_scoped_release r;
               ok(hiqsim_measure_qubits(h_, ids.data(), ids.size(), out.data()));
          }
          return std::vector<bool>(out.begin(), out.end());
     }
     void apply_controlled_gate(
         py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> m,
         const std::vector<int64_t>& ids, const std::vector<int64_t>& ctrls)
     {
          if (m.ndim() != 2 || m.shape(0) != m.shape(1))
               throw std::runtime_error("apply_controlled_gate(): matrix must be square");
          const auto* d = reinterpret_cast<const double*>(m.data());
          const size_t dim = static_cast<size_t>(m.shape(0));
          ok(hiqsim_apply_controlled_gate(h_, d, dim, ids.data(), ids.size(), ctrls.data(),
                                          ctrls.size()));
     }
     void emulate_math(const py::object&, const py::object&, const py::object&)
     {
          // SimulatorMPI.hpp:217-225.
          throw std::runtime_error("SimulatorMPI::emulate_math() is not supported");
     }
     std::complex<double> get_amplitude(const std::vector<bool>& bits,
                                        const std::vector<int64_t>& ids)
     {
          auto b = to_bits(bits);
          double out[2];
          {
               py::gil_scoped_release r;
               ok(hiqsim_get_amplitude(h_, b.data(), b.size(), ids.data(), ids.size(), out));
          }
          return {out[0], out[1]};
     }
     double get_probability(const std::vector<bool>& bits, const std::vector<int64_t>& ids)
     {
          auto b = to_bits(bits);
          double out = 0;
          {
               py::gil_scoped_release r;
               ok(hiqsim_get_probability(h_, b.data(), b.size(), ids.data(), ids.size(), &out));
          }
          return out;
     }
     void run()
     {
          py::gil_scoped_release r;
          ok(hiqsim_run(h_));
     }
     void set_pass_mode(int mode)
     {
          py::gil_scoped_release r;
          ok(hiqsim_set_pass_mode(h_, mode));
     }
     py::tuple pass_mode()
     {
          int mode = 0;
          size_t queued = 0;
          ok(hiqsim_get_pass_mode(h_, &mode, &queued));
          return py::make_tuple(mode, queued);
     }
     double entropy()
     {
          double e = 0;
          py::gil_scoped_release r;
          ok(hiqsim_entropy(h_, &e));
          return e;
     }
     py::array_t<std::complex<double>> get_state()
     {
          uint64_t n = 0;
          ok(hiqsim_state_size(h_, &n));
          py::array_t<std::complex<double>> a(static_cast<py::ssize_t>(n));
          auto* p = reinterpret_cast<double*>(a.mutable_data());
          {
               py::gil_scoped_release r;
               ok(hiqsim_get_state(h_, p, n));
          }
          return a;
     }
     // The state without a copy, as a DLPack capsule ("dltensor") for
     // torch.from_dlpack: 2^n complex128/complex64 amplitudes on the GPU.
     // The view pins the buffer (hiqsim_pin_device_state) until the tensor
     // is freed: permutations and Pauli sums then write their results back
     // into it, and the memory stays allocated; the view follows the state
     // until the qubit count changes.
     static py::capsule state_dlpack(py::object self_obj)
     {
          PySim& self = self_obj.cast<PySim&>();
          void* ptr = nullptr;
          uint64_t n = 0;
          int prec = 0, dev = 0;
          {
               py::gil_scoped_release r;
               ok(hiqsim_pin_device_state(self.h_, &ptr, &n, &prec, &dev));
          }
          auto* v = new StateView();
          v->handle = self.h_;
          v->shape[0] = static_cast<int64_t>(n);
          v->t.dl_tensor.data = ptr;
          v->t.dl_tensor.device = DLDevice{kDLROCM, dev};
          v->t.dl_tensor.ndim = 1;
          v->t.dl_tensor.dtype = DLDataType{kDLComplex, static_cast<uint8_t>(prec == HIQSIM_C128 ? 128 : 64), 1};
          v->t.dl_tensor.shape = v->shape;
          v->t.dl_tensor.strides = nullptr;
          v->t.dl_tensor.byte_offset = 0;
          v->t.manager_ctx = v;
          v->t.deleter = state_view_deleter;
          v->owner = self_obj.ptr();
          Py_INCREF(v->owner);
          return py::capsule(&v->t, "dltensor", [](PyObject* cap) {
               // an unconsumed capsule still owns the view
               if (PyCapsule_IsValid(cap, "dltensor")) {
                    auto* t = static_cast<DLManagedTensor*>(PyCapsule_GetPointer(cap, "dltensor"));
                    t->deleter(t);
               }
          });
     }
     void save_state(const std::string& path)
     {
          py::gil_scoped_release r;
          ok(hiqsim_save_state(h_, path.c_str()));
     }
     void load_state(const std::string& path)
     {
          py::gil_scoped_release r;
          ok(hiqsim_load_state(h_, path.c_str()));
     }
     // SimulatorMPI::cheat_local (SimulatorMPI.cpp:543-559): id -> bit
     // position and a copy of the state.
     py::tuple cheat_local()
     {
          std::map<int64_t, int64_t> id2pos;
          auto ids = qubit_ids();
          for (size_t p = 0; p < ids.size(); ++p) id2pos[ids[p]] = static_cast<int64_t>(p);
          return py::make_tuple(id2pos, get_state());
     }
     void set_state(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> v)
     {
          const auto* p = reinterpret_cast<const double*>(v.data());
          const uint64_t n = static_cast<uint64_t>(v.size());
          py::gil_scoped_release r;
          ok(hiqsim_set_state(h_, p, n));
     }
     // set_wavefunction(wf, ids) (_simulator_mpi.py:281-306): the simulator
     // adopts the ordering of `ids` for the given amplitudes.
     void set_wavefunction(
         py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> wf,
         const std::vector<int64_t>& ids)
     {
          auto cur = qubit_ids();
          if (ids.size() != cur.size())
               throw std::runtime_error(
                   "set_wavefunction(): the qureg must contain all allocated qubits");
          if (static_cast<uint64_t>(wf.size()) != (uint64_t(1) << ids.size()))
               throw std::runtime_error("set_wavefunction(): wrong number of amplitudes");
          std::vector<int64_t> a(ids), b(cur);
          std::sort(a.begin(), a.end());
          std::sort(b.begin(), b.end());
          if (a != b)
               throw std::runtime_error(
                   "set_wavefunction(): the qureg must be a permutation of the allocated qubits");
          ok(hiqsim_set_qubits_perm(h_, ids.data(), ids.size()));
          set_state(wf);
     }
     void collapse_wavefunction(const std::vector<int64_t>& ids, const std::vector<bool>& values)
     {
          auto v = to_bits(values);
          py::gil_scoped_release r;
          ok(hiqsim_collapse_wavefunction(h_, ids.data(), ids.size(), v.data(), v.size()));
     }
     double get_expectation_value(const py::list& terms, const std::vector<int64_t>& ids)
     {
          std::vector<uint64_t> X, Z;
          std::vector<double> C;
          parse_terms(terms, X, Z, C);
          double out[2] = {0, 0};
          {
               py::gil_scoped_release r;
               ok(hiqsim_expectation_pauli(h_, X.data(), Z.data(), C.data(), X.size(), ids.data(),
                                           ids.size(), out));
          }
          return out[0];
     }
     std::complex<double> expectation_complex(const py::list& terms,
                                              const std::vector<int64_t>& ids)
     {
          std::vector<uint64_t> X, Z;
          std::vector<double> C;
          parse_terms(terms, X, Z, C);
          double out[2] = {0, 0};
          {
               py::gil_scoped_release r;
               ok(hiqsim_expectation_pauli(h_, X.data(), Z.data(), C.data(), X.size(), ids.data(),
                                           ids.size(), out));
          }
          return {out[0], out[1]};
     }
     void apply_qubit_operator(const py::list& terms, const std::vector<int64_t>& ids)
     {
          std::vector<uint64_t> X, Z;
          std::vector<double> C;
          parse_terms(terms, X, Z, C);
          py::gil_scoped_release r;
          ok(hiqsim_apply_pauli_sum(h_, X.data(), Z.data(), C.data(), X.size(), ids.data(),
                                    ids.size()));
     }
     // _simulator_mpi.py:469-475: the call the reference engine makes for a
     // TimeEvolution gate (terms as for apply_qubit_operator).
     void emulate_time_evolution(const py::list& terms, double time, const std::vector<int64_t>& ids,
                                 const std::vector<int64_t>& ctrls)
     {
          std::vector<uint64_t> X, Z;
          std::vector<double> C;
          parse_terms(terms, X, Z, C);
          py::gil_scoped_release r;
          ok(hiqsim_emulate_time_evolution(h_, X.data(), Z.data(), C.data(), X.size(), time, ids.data(),
                                           ids.size(), ctrls.data(), ctrls.size()));
     }
     void apply_matrix(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> m,
                       const std::vector<uint32_t>& pos, uint64_t ctrlmask, bool diag)
     {
          if (m.ndim() != 2 || m.shape(0) != (py::ssize_t(1) << pos.size()))
               throw std::runtime_error("apply_matrix(): matrix does not match positions");
          const auto* d = reinterpret_cast<const double*>(m.data());
          py::gil_scoped_release r;
          ok(hiqsim_apply_matrix(h_, d, static_cast<int>(pos.size()), pos.data(), ctrlmask,
                                 diag ? 1 : 0));
     }
     void snapshot()
     {
          py::gil_scoped_release r;
          ok(hiqsim_snapshot_state(h_));
     }
     void restore()
     {
          py::gil_scoped_release r;
          ok(hiqsim_restore_state(h_));
     }
     py::dict compare_snapshot()
     {
          double out[5];
          {
               py::gil_scoped_release r;
               ok(hiqsim_compare_snapshot(h_, out));
          }
          py::dict d;
          d["overlap"] = std::complex<double>(out[0], out[1]);
          d["ref_norm2"] = out[2];
          d["max_abs_diff"] = out[3];
          d["state_norm2"] = out[4];
          // |<ref|psi>|^2 / (<ref|ref> <psi|psi>)
          d["fidelity"] = (out[0] * out[0] + out[1] * out[1]) / (out[2] * out[4]);
          return d;
     }
     void init_random_state(uint64_t seed)
     {
          py::gil_scoped_release r;
          ok(hiqsim_init_random_state(h_, seed));
     }
     int apply_tiled(const py::list& gates, int op_width)
     {
          const GateCsr g = gate_csr(gates);
          int np = 0;
          py::gil_scoped_release r;
          ok(hiqsim_apply_tiled(h_, g.size(), g.mats.data(), g.tid.data(), g.toff.data(), g.cid.data(),
                                g.coff.data(), op_width, &np));
          return np;
     }
     py::array_t<uint64_t> sample(uint32_t shots, uint64_t seed)
     {
          py::array_t<uint64_t> a(static_cast<py::ssize_t>(shots));
          uint64_t* d = a.mutable_data();
          {
               py::gil_scoped_release r;
               ok(hiqsim_sample(h_, shots, seed, d));
          }
          return a;
     }
     void synchronize()
     {
          py::gil_scoped_release r;
          ok(hiqsim_synchronize(h_));
     }
     void set_timing(bool on) { ok(hiqsim_set_timing(h_, on ? 1 : 0)); }
     py::tuple kernel_time(int cls)
     {
          uint64_t c = 0;
          double ms = 0;
          {
               py::gil_scoped_release r;
               ok(hiqsim_kernel_time(h_, cls, &c, &ms));
          }
          return py::make_tuple(c, ms);
     }
     uint64_t state_size() const
     {
          uint64_t n = 0;
          ok(hiqsim_state_size(h_, &n));
          return n;
     }

private:
     hiqsim_t* h_ = nullptr;
};

struct CsrGates {
     std::vector<int64_t> gid, cid;
     std::vector<size_t> goff{0}, coff{0};
     std::vector<uint8_t> diag;
     CsrGates(const std::vector<std::vector<int64_t>>& gate,
              const std::vector<std::vector<int64_t>>& ctrl, const std::vector<bool>& d)
     {
          if (gate.size() != ctrl.size() || gate.size() != d.size())
               throw std::runtime_error("ctor(): gate, gate_ctrl and gate_diag differ in length");
          for (size_t i = 0; i < gate.size(); ++i) {
               gid.insert(gid.end(), gate[i].begin(), gate[i].end());
               cid.insert(cid.end(), ctrl[i].begin(), ctrl[i].end());
               goff.push_back(gid.size());
               coff.push_back(cid.size());
               diag.push_back(d[i] ? 1 : 0);
          }
     }
};

// _sched_cpp.ClusterScheduler (_sched_cpp.cpp:36-39).
class PyClusterScheduler {
public:
     PyClusterScheduler(std::vector<std::vector<int64_t>> gate,
                        std::vector<std::vector<int64_t>> ctrl, std::vector<bool> diag,
                        std::vector<int64_t> locals, std::vector<int64_t> globals, int size)
         : g_(gate, ctrl, diag), locals_(std::move(locals)), globals_(std::move(globals)), size_(size)
     {}
     std::vector<int> schedule()
     {
          std::vector<int32_t> out(g_.diag.size() + 1);
          size_t n = 0;
          {
               py::gil_scoped_release r;
               ok(hiqsim_schedule_cluster(g_.gid.data(), g_.goff.data(), g_.cid.data(),
                                          g_.coff.data(), g_.diag.data(), g_.diag.size(),
                                          locals_.data(), locals_.size(), globals_.data(),
                                          globals_.size(), size_, out.data(), &n));
          }
          return std::vector<int>(out.begin(), out.begin() + static_cast<long>(n));
     }

private:
     CsrGates g_;
     std::vector<int64_t> locals_, globals_;
     int size_;
};

// _sched_cpp.SwapScheduler (_sched_cpp.cpp:31-35).
class PySwapScheduler {
public:
     PySwapScheduler(std::vector<std::vector<int64_t>> gate, std::vector<std::vector<int64_t>> ctrl,
                     std::vector<bool> diag, int64_t num_splits, int num_locals, bool fuse)
         : g_(gate, ctrl, diag), splits_(num_splits), locals_(num_locals), fuse_(fuse)
     {}
     std::vector<int64_t> schedule()
     {
          std::vector<int64_t> out(64);
          size_t n = 0;
          {
               py::gil_scoped_release r;
               ok(hiqsim_schedule_swap(g_.gid.data(), g_.goff.data(), g_.cid.data(), g_.coff.data(),
                                       g_.diag.data(), g_.diag.size(), splits_, locals_,
                                       fuse_ ? 1 : 0, out.data(), &n));
          }
          out.resize(n);
          return out;
     }

private:
     CsrGates g_;
     int64_t splits_;
     int locals_;
     bool fuse_;
};

// Many independent circuits run together (hiqsim_batch_*).
class PyBatch {
public:
     PyBatch(const std::string& precision, int device)
     {
          int p;
          if (precision == "c128" || precision == "complex128") p = HIQSIM_C128;
          else if (precision == "c64" || precision == "complex64") p = HIQSIM_C64;
          else throw std::runtime_error("precision must be 'c128' or 'c64'");
          ok(hiqsim_batch_create(p, device, &h_));
     }
     ~PyBatch()
     {
          if (h_) hiqsim_batch_destroy(h_);
     }
     // gates: [(matrix, targets, controls), ...]
     int64_t add_circuit(int n, const py::list& gates)
     {
          const GateCsr g = gate_csr(gates);
          int64_t idx = -1;
          ok(hiqsim_batch_add_circuit(h_, n, g.size(), g.mats.data(), g.tid.data(), g.toff.data(),
                                      g.cid.data(), g.coff.data(), &idx));
          return idx;
     }
     int64_t add_copies(int64_t index, int64_t count)
     {
          int64_t first = -1;
          ok(hiqsim_batch_add_copies(h_, index, count, &first));
          return first;
     }
     void set_plan(const std::map<int, int>& k_by_n, int wg_max_qubits, int tile_min_qubits,
                   int prog_min_qubits)
     {
          std::vector<int32_t> v;
          for (auto& kv: k_by_n) {
               if (kv.first < 0 || kv.first > 64) throw std::runtime_error("qubit count out of range");
               if (v.size() <= static_cast<size_t>(kv.first)) v.resize(kv.first + 1, 0);
               v[kv.first] = kv.second;
          }
          ok(hiqsim_batch_set_plan(h_, v.data(), v.size(), wg_max_qubits, tile_min_qubits,
                                   prog_min_qubits));
     }
     void prepare()
     {
          py::gil_scoped_release r;
          ok(hiqsim_batch_prepare(h_));
     }
     void build()
     {
          py::gil_scoped_release r;
          ok(hiqsim_batch_build(h_));
     }
     double run()
     {
          double ms = 0;
          py::gil_scoped_release r;
          ok(hiqsim_batch_run(h_, &ms));
          return ms;
     }
     int64_t size() const
     {
          int64_t n = 0;
          ok(hiqsim_batch_num_circuits(h_, &n));
          return n;
     }
     int num_qubits(int64_t i) const
     {
          int n = 0;
          ok(hiqsim_batch_num_qubits(h_, i, &n));
          return n;
     }
     py::array_t<std::complex<double>> get_state(int64_t i) const
     {
          const int n = num_qubits(i);
          py::array_t<std::complex<double>> a(static_cast<py::ssize_t>(1) << n);
          double* d = reinterpret_cast<double*>(a.mutable_data());
          {
               py::gil_scoped_release r;
               ok(hiqsim_batch_get_state(h_, i, d));
          }
          return a;
     }
     py::dict stats() const
     {
          uint64_t v[14] = {0};
          ok(hiqsim_batch_stats(h_, v, 14));
          static const char* names[14] = {"circuits",      "passes",         "wg_circuits",
                                          "lds_circuits",  "layer_launches", "arena_bytes",
                                          "matrix_bytes",  "amp_passes",     "wg_grid",
                                          "kmax",          "padded_launches", "tile_passes",
                                          "tile_ops",      "tile_mfma_flops"};
          py::dict d;
          for (int i = 0; i < 14; ++i) d[names[i]] = v[i];
          return d;
     }

private:
     hiqsim_batch_t* h_ = nullptr;
};

// Request loop (hiqsim_server_*).
class PyServer {
public:
     PyServer(const std::string& precision, int device, uint64_t state_bytes, uint64_t matrix_bytes,
              int passes_per_step, int max_active)
     {
          int p;
          if (precision == "c128" || precision == "complex128") p = HIQSIM_C128;
          else if (precision == "c64" || precision == "complex64") p = HIQSIM_C64;
          else throw std::runtime_error("precision must be 'c128' or 'c64'");
          ok(hiqsim_server_create(p, device, state_bytes, matrix_bytes, passes_per_step, max_active, &h_));
          max_active_ = static_cast<size_t>(max_active);
     }
     ~PyServer()
     {
          if (h_) hiqsim_server_destroy(h_);
     }
     void set_widths(const std::map<int, int>& k_by_n)
     {
          std::vector<int32_t> v;
          for (auto& kv: k_by_n) {
               if (kv.first < 0 || kv.first > 64) throw std::runtime_error("qubit count out of range");
               if (v.size() <= static_cast<size_t>(kv.first)) v.resize(kv.first + 1, 0);
               v[kv.first] = kv.second;
          }
          ok(hiqsim_server_set_widths(h_, v.data(), v.size()));
     }
     void set_prog_min_qubits(int q) { ok(hiqsim_server_set_prog_min_qubits(h_, q)); }
     void set_heavy_qubits(int q) { ok(hiqsim_server_set_heavy_qubits(h_, q)); }
     int64_t submit_samples(int n, const py::list& gates, uint32_t shots, uint64_t seed)
     {
          const GateCsr g = gate_csr(gates);
          int64_t id = -1;
          ok(hiqsim_server_submit(h_, n, g.size(), g.mats.data(), g.tid.data(), g.toff.data(),
                                  g.cid.data(), g.coff.data(), 0, shots, seed, nullptr, nullptr, 0, &id));
          return id;
     }
     int64_t submit_expectation(int n, const py::list& gates, const std::vector<uint64_t>& xmask,
                                const std::vector<uint64_t>& zmask)
     {
          const GateCsr g = gate_csr(gates);
          if (xmask.size() != zmask.size()) throw std::runtime_error("xmask and zmask differ in length");
          int64_t id = -1;
          ok(hiqsim_server_submit(h_, n, g.size(), g.mats.data(), g.tid.data(), g.toff.data(),
                                  g.cid.data(), g.coff.data(), 1, 0, 0, xmask.data(), zmask.data(),
                                  xmask.size(), &id));
          return id;
     }
     int64_t submit_state(int n, const py::list& gates)
     {
          const GateCsr g = gate_csr(gates);
          int64_t id = -1;
          ok(hiqsim_server_submit(h_, n, g.size(), g.mats.data(), g.tid.data(), g.toff.data(),
                                  g.cid.data(), g.coff.data(), 2, 0, 0, nullptr, nullptr, 0, &id));
          return id;
     }
     std::vector<int64_t> step(bool block)
     {
          std::vector<int64_t> done(max_active_ + 16);
          size_t n = 0;
          {
               py::gil_scoped_release r;
               ok(hiqsim_server_step(h_, block ? 1 : 0, done.data(), done.size(), &n));
          }
          done.resize(std::min(n, done.size()));
          return done;
     }
     bool idle() const
     {
          int v = 0;
          ok(hiqsim_server_idle(h_, &v));
          return v != 0;
     }
     bool done(int64_t id) const
     {
          int v = 0;
          ok(hiqsim_server_done(h_, id, &v));
          return v != 0;
     }
     py::array_t<uint64_t> samples(int64_t id) const
     {
          size_t n = 0;
          ok(hiqsim_server_samples(h_, id, nullptr, 0, &n));
          py::array_t<uint64_t> a(static_cast<py::ssize_t>(n));
          ok(hiqsim_server_samples(h_, id, a.mutable_data(), n, &n));
          return a;
     }
     py::array_t<std::complex<double>> expectations(int64_t id) const
     {
          size_t n = 0;
          ok(hiqsim_server_expectations(h_, id, nullptr, 0, &n));
          py::array_t<std::complex<double>> a(static_cast<py::ssize_t>(n));
          ok(hiqsim_server_expectations(h_, id, reinterpret_cast<double*>(a.mutable_data()), n, &n));
          return a;
     }
     py::array_t<std::complex<double>> state(int64_t id) const
     {
          uint64_t n = 0;
          ok(hiqsim_server_state(h_, id, nullptr, 0, &n));
          py::array_t<std::complex<double>> a(static_cast<py::ssize_t>(n));
          ok(hiqsim_server_state(h_, id, reinterpret_cast<double*>(a.mutable_data()), n, &n));
          return a;
     }
     py::tuple times(int64_t id) const
     {
          double t[3];
          ok(hiqsim_server_times(h_, id, t));
          return py::make_tuple(t[0], t[1], t[2]);
     }
     void forget(int64_t id) { ok(hiqsim_server_forget(h_, id)); }
     py::dict stats() const
     {
          uint64_t v[6] = {0};
          ok(hiqsim_server_stats(h_, v, 6));
          static const char* names[6] = {"submitted", "completed", "steps", "launches", "peak_active",
                                         "peak_state_bytes"};
          py::dict d;
          for (int i = 0; i < 6; ++i) d[names[i]] = v[i];
          return d;
     }

private:
     hiqsim_server_t* h_ = nullptr;
     size_t max_active_ = 0;
};

}  // namespace

PYBIND11_MODULE(_cppsim_amd, m)
{
     m.doc() = "MI355X state-vector engine (drop-in for HiQsimulator's _cppsim_mpi)";
     m.def("device_count", [] {
          int n = 0;
          ok(hiqsim_device_count(&n));
          return n;
     });
     py::class_<PySim>(m, "SimulatorMPI")
         .def(py::init<uint64_t, int64_t, int64_t, const std::string&, int>(), py::arg("seed"),
              py::arg("max_local"), py::arg("max_cluster_size"), py::arg("precision") = "c128",
              py::arg("device") = 0)
         .def("get_qubits_ids", &PySim::qubit_ids)
         .def("get_local_qubits_ids", &PySim::qubit_ids)
         .def("get_global_qubits_ids", &PySim::global_ids)
         .def("set_qubits_perm", &PySim::set_qubits_perm)
         .def("swap_qubits", &PySim::swap_qubits)
         .def("permute_qubits", &PySim::permute_qubits)
         .def("allocate_qureg", &PySim::allocate_qureg)
         .def("allocate_qubit", &PySim::allocate_qubit)
         .def("deallocate_qubit", &PySim::deallocate_qubit)
         .def("measure_qubits", &PySim::measure_qubits)
         .def("apply_controlled_gate", &PySim::apply_controlled_gate)
         .def("emulate_math", &PySim::emulate_math)
         .def("get_amplitude", &PySim::get_amplitude)
         .def("get_probability", &PySim::get_probability)
         .def("run", &PySim::run)
         .def("set_pass_mode", &PySim::set_pass_mode)
         .def("pass_mode", &PySim::pass_mode)
         .def("entropy", &PySim::entropy)
         .def("cheat_local", &PySim::cheat_local)
         .def("collapse_wavefunction", &PySim::collapse_wavefunction)
         .def("get_state", &PySim::get_state)
         .def("state_dlpack", &PySim::state_dlpack)
         .def("save_state", &PySim::save_state)
         .def("load_state", &PySim::load_state)
         .def("set_state", &PySim::set_state)
         .def("set_wavefunction", &PySim::set_wavefunction)
         .def("get_expectation_value", &PySim::get_expectation_value)
         .def("expectation_complex", &PySim::expectation_complex)
         .def("apply_qubit_operator", &PySim::apply_qubit_operator)
         .def("emulate_time_evolution", &PySim::emulate_time_evolution)
         .def("apply_matrix", &PySim::apply_matrix, py::arg("matrix"), py::arg("pos"),
              py::arg("ctrlmask") = 0, py::arg("diag") = false)
         .def("snapshot", &PySim::snapshot)
         .def("compare_snapshot", &PySim::compare_snapshot)
         .def("init_random_state", &PySim::init_random_state)
         .def("sample", &PySim::sample, py::arg("shots"), py::arg("seed"))
         .def("apply_tiled", &PySim::apply_tiled, py::arg("gates"), py::arg("op_width") = 4)
         .def("restore", &PySim::restore)
         .def("synchronize", &PySim::synchronize)
         .def("set_timing", &PySim::set_timing)
         .def("kernel_time", &PySim::kernel_time)
         .def("state_size", &PySim::state_size);
     m.def("set_tuning", [](const std::string& name, int64_t v) { ok(hiqsim_set_tuning(name.c_str(), v)); });
     m.def("get_tuning", [](const std::string& name) {
          int64_t v = 0;
          ok(hiqsim_get_tuning(name.c_str(), &v));
          return v;
     });
     m.def("plan_tile_passes",
           [](const std::vector<std::vector<int64_t>>& gate,
              const std::vector<std::vector<int64_t>>& ctrl, int n_bits, int op_width) {
                CsrGates g(gate, ctrl, std::vector<bool>(gate.size(), false));
                std::vector<int32_t> pass_of(gate.size()), op_of(gate.size()), hpos(6 * gate.size() + 6);
                size_t np = 0;
                {
                     py::gil_scoped_release r;
                     ok(hiqsim_plan_tile_passes(g.gid.data(), g.goff.data(), g.cid.data(), g.coff.data(),
                                                gate.size(), n_bits, op_width, pass_of.data(), op_of.data(),
                                                hpos.data(), &np));
                }
                hpos.resize(6 * np);
                return py::make_tuple(pass_of, op_of, hpos);
           },
           py::arg("gate"), py::arg("gate_ctrl"), py::arg("n_bits"), py::arg("op_width") = 4);
     m.def("plan_clusters",
           [](const std::vector<std::vector<int64_t>>& gate,
              const std::vector<std::vector<int64_t>>& ctrl, int cluster_size, int mode) {
                CsrGates g(gate, ctrl, std::vector<bool>(gate.size(), false));
                std::vector<int32_t> out(gate.size());
                {
                     py::gil_scoped_release r;
                     ok(hiqsim_plan_clusters(g.gid.data(), g.goff.data(), g.cid.data(),
                                             g.coff.data(), gate.size(), cluster_size, mode,
                                             out.data()));
                }
                return out;
           },
           py::arg("gate"), py::arg("gate_ctrl"), py::arg("cluster_size"), py::arg("mode") = 0);
     py::class_<PyServer>(m, "Server")
         .def(py::init<const std::string&, int, uint64_t, uint64_t, int, int>(),
              py::arg("precision") = "c64", py::arg("device") = 0,
              py::arg("state_bytes") = uint64_t(32) << 30, py::arg("matrix_bytes") = uint64_t(2) << 30,
              py::arg("passes_per_step") = 8, py::arg("max_active") = 1024)
         .def("set_widths", &PyServer::set_widths)
         .def("set_prog_min_qubits", &PyServer::set_prog_min_qubits)
         .def("set_heavy_qubits", &PyServer::set_heavy_qubits)
         .def("submit_samples", &PyServer::submit_samples, py::arg("n_qubits"), py::arg("gates"),
              py::arg("shots") = 1024, py::arg("seed") = 0)
         .def("submit_expectation", &PyServer::submit_expectation, py::arg("n_qubits"),
              py::arg("gates"), py::arg("xmask"), py::arg("zmask"))
         .def("submit_state", &PyServer::submit_state, py::arg("n_qubits"), py::arg("gates"))
         .def("step", &PyServer::step, py::arg("block") = false)
         .def("idle", &PyServer::idle)
         .def("done", &PyServer::done)
         .def("samples", &PyServer::samples)
         .def("expectations", &PyServer::expectations)
         .def("state", &PyServer::state)
         .def("times", &PyServer::times)
         .def("forget", &PyServer::forget)
         .def("stats", &PyServer::stats);
     py::class_<PyBatch>(m, "Batch")
         .def(py::init<const std::string&, int>(), py::arg("precision") = "c128",
              py::arg("device") = 0)
         .def("add_circuit", &PyBatch::add_circuit, py::arg("n_qubits"), py::arg("gates"))
         .def("add_copies", &PyBatch::add_copies, py::arg("index"), py::arg("count"))
         .def("set_plan", &PyBatch::set_plan, py::arg("k_by_n"), py::arg("wg_max_qubits") = 16,
              py::arg("tile_min_qubits") = 0, py::arg("prog_min_qubits") = 31)
         .def("prepare", &PyBatch::prepare)
         .def("build", &PyBatch::build)
         .def("run", &PyBatch::run)
         .def("__len__", &PyBatch::size)
         .def("num_qubits", &PyBatch::num_qubits)
         .def("get_state", &PyBatch::get_state)
         .def("stats", &PyBatch::stats);
     py::class_<PySwapScheduler>(m, "SwapScheduler")
         .def(py::init<std::vector<std::vector<int64_t>>, std::vector<std::vector<int64_t>>,
                       std::vector<bool>, int64_t, int, bool>())
         .def("ScheduleSwap", &PySwapScheduler::schedule);
     py::class_<PyClusterScheduler>(m, "ClusterScheduler")
         .def(py::init<std::vector<std::vector<int64_t>>, std::vector<std::vector<int64_t>>,
                       std::vector<bool>, std::vector<int64_t>, std::vector<int64_t>, int>())
         .def("ScheduleCluster", &PyClusterScheduler::schedule);
}
