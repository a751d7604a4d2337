// CPU ORACLE — test infrastructure only.
//
// A plain FP64 restatement of HiQsimulator's full-amplitude SimulatorMPI at
// MPI world size 1 (no global qubits).  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this library, and only as the
// checker: the product path (hiqsimulator_amd) never links or calls it.
//
// Every function names the reference file:line it restates (paths relative
// to the reference checkout root).  Build: oracle/Makefile, flags
// -O2 -ffp-contract=off (no fast-math), so results are reproducible.
//
// Parity pins (see tests/test_oracle_*.py):
//   dense / diagonal passes  -> oracle/_ref (the reference's own kernel
//                               headers compiled unmodified) + golden fixtures
//   fusion / control folding -> brute-force sequential application, the
//                               reference's kron known answer
//   measurement RNG          -> independent numpy MT19937 + generate_canonical
//                               restatement, the C++ standard's 10000th value
//   probability / amplitude / collapse / dealloc / entropy -> the reference
//                               test-suite known answers and closed forms
//   permutation              -> closed form out[sigma(i)] = in[i], brute force
//   Pauli expectation        -> dense kron matrices (numpy) + commented-out
//                               reference known answers
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using Index = int64_t;
using Complex = std::complex<double>;
using Row = std::vector<Complex>;
using Matrix = std::vector<Row>;
constexpr size_t kNotFound = static_cast<size_t>(-1);
constexpr double kMaxFloatError = 1e-12;  // SimulatorMPI.cpp:75

thread_local std::string g_last_error;

// Complex product written out (no __muldc3 special cases, no contraction).
inline Complex cmul(Complex a, Complex b)
{
     return Complex(a.real() * b.real() - a.imag() * b.imag(),
                    a.real() * b.imag() + a.imag() * b.real());
}

inline double cnorm(Complex a)
{
     return a.real() * a.real() + a.imag() * a.imag();
}

// ---------------------------------------------------------------------------
// Kernels.  kernels/nointrin/kernel{1..5}.hpp: bases I enumerated by the
// descending-stride loop nest (target bits zero), skipped unless
// (I & ctrlmask) == ctrlmask; v[j] = psi[I + sum_{l in j} d_l] with
// d_l = 1 << pos[l]; psi[I + off_i] = sum_j m[i][j] v[j].
// ---------------------------------------------------------------------------
void dense_pass(Complex* psi, uint64_t n, const Matrix& m,
                const std::vector<uint32_t>& pos, uint64_t ctrlmask)
{
     const unsigned k = static_cast<unsigned>(pos.size());
     const uint64_t dim = 1ull << k;
     std::vector<uint64_t> off(dim, 0);
     for (uint64_t j = 0; j < dim; ++j)
          for (unsigned l = 0; l < k; ++l)
               if (j >> l & 1) off[j] += 1ull << pos[l];
     uint64_t tmask = 0;
     for (auto p: pos) tmask |= 1ull << p;
     std::vector<Complex> v(dim);
     // The nested loops of nointrin::kernel visit exactly the indices with
     // all target bits zero, in increasing order.
     for (uint64_t I = 0; I < n; ++I) {
          if (I & tmask) continue;
          if ((I & ctrlmask) != ctrlmask) continue;
          for (uint64_t j = 0; j < dim; ++j) v[j] = psi[I + off[j]];
          for (uint64_t i = 0; i < dim; ++i) {
               Complex acc(0.0, 0.0);
               for (uint64_t j = 0; j < dim; ++j) acc += cmul(m[i][j], v[j]);
               psi[I + off[i]] = acc;
          }
     }
}

// kernels/intrin/kernels_diag.hpp:34-144 (nointrin/kernels_diag.hpp:19-84):
// every index i with (i & cmask) == cmask is scaled by m[d][d], d built from
// the target bits of i.
void diag_pass(Complex* psi, uint64_t n, const Matrix& m,
               const std::vector<uint32_t>& pos, uint64_t cmask)
{
     for (uint64_t i = 0; i < n; ++i) {
          if ((i & cmask) != cmask) continue;
          uint64_t d = 0;
          for (size_t l = 0; l < pos.size(); ++l) d |= ((i >> pos[l]) & 1ull) << l;
          psi[i] = cmul(psi[i], m[d][d]);
     }
}

// kernels/nointrin/kernels_diag.hpp:8-17 (kernelK_diag1): global factor.
void scale_pass(Complex* psi, uint64_t n, Complex f)
{
     for (uint64_t i = 0; i < n; ++i) psi[i] = cmul(psi[i], f);
}

// funcs.hpp:325-347 — exact != 0.0 test off the diagonal.
bool is_diag(const Matrix& m)
{
     for (size_t i = 0; i < m.size(); ++i)
          for (size_t j = 0; j < m[i].size(); ++j)
               if (i != j && m[i][j] != 0.0) return false;
     return true;
}

// ---------------------------------------------------------------------------
// Fusion: fusion_mpi.hpp:64-234.
// ---------------------------------------------------------------------------
struct Item {
     Matrix mat;
     std::vector<Index> idx;
     bool diag;
};

struct Fusion {
     std::set<Index> set_;
     std::vector<Item> items_;
     std::set<Index> ctrl_set_;
     Complex factor = 1.0;

     // fusion_mpi.hpp:167-187: controls become the highest matrix bits,
     // newmatrix = diag(I_{Offset}, U) with Offset = F*d - d.
     static void add_controls(Matrix& matrix, std::vector<Index>& ids,
                              const std::vector<Index>& new_ctrls)
     {
          ids.insert(ids.end(), new_ctrls.begin(), new_ctrls.end());
          size_t F = size_t(1) << new_ctrls.size();
          size_t d = matrix.size();
          Matrix nm(F * d, Row(F * d, Complex(0.0, 0.0)));
          size_t offset = F * d - d;
          for (size_t i = 0; i < offset; ++i) nm[i][i] = 1.0;
          for (size_t i = 0; i < d; ++i)
               for (size_t j = 0; j < d; ++j) nm[offset + i][offset + j] = matrix[i][j];
          matrix = std::move(nm);
     }

     // fusion_mpi.hpp:190-226.
     void handle_controls(Matrix& matrix, std::vector<Index>& ids,
                          const std::vector<Index>& ctrls)
     {
          auto unhandled = ctrl_set_;
          for (auto c: ctrls) {
               if (ctrl_set_.count(c) == 0) {
                    if (!items_.empty()) {
                         add_controls(matrix, ids, {c});
                         set_.insert(c);
                    }
                    else {
                         ctrl_set_.insert(c);
                    }
               }
               else {
                    unhandled.erase(c);
               }
          }
          if (!unhandled.empty()) {
               std::vector<Index> nc(unhandled.begin(), unhandled.end());
               for (auto c: nc) {
                    ctrl_set_.erase(c);
                    set_.insert(c);
               }
               for (auto& it: items_) add_controls(it.mat, it.idx, nc);
          }
     }

     // fusion_mpi.hpp:85-99.
     void insert(Matrix matrix, bool diag, std::vector<Index> ids,
                 const std::vector<Index>& ctrls)
     {
          if (matrix.size() == 1) {
               factor = factor * matrix[0][0];
               return;
          }
          for (auto i: ids) set_.insert(i);
          handle_controls(matrix, ids, ctrls);
          items_.push_back(Item{matrix, ids, diag});
     }

     // fusion_mpi.hpp:101-165: M = G_last ... G_1 * (factor * I) over the
     // ascending id list; item bit l <-> position of idx[l] in that list.
     void perform(Matrix& M, std::vector<Index>& ids, std::vector<Index>& ctrls,
                  bool& diag)
     {
          ids.assign(set_.begin(), set_.end());
          size_t N = ids.size();
          size_t dim = size_t(1) << N;
          M.assign(dim, Row(dim, Complex(0.0, 0.0)));
          for (size_t i = 0; i < dim; ++i) M[i][i] = factor;
          diag = true;
          for (auto& item: items_) {
               if (!item.diag) diag = false;
               std::vector<size_t> idx2mat(item.idx.size());
               for (size_t i = 0; i < item.idx.size(); ++i)
                    idx2mat[i] = static_cast<size_t>(
                        std::lower_bound(ids.begin(), ids.end(), item.idx[i]) - ids.begin());
               const size_t kk = item.idx.size();
               std::vector<Complex> oldcol(dim);
               for (size_t col = 0; col < dim; ++col) {
                    for (size_t i = 0; i < dim; ++i) oldcol[i] = M[i][col];
                    for (size_t i = 0; i < dim; ++i) {
                         size_t li = 0;
                         for (size_t l = 0; l < kk; ++l) li |= ((i >> idx2mat[l]) & 1) << l;
                         Complex res(0.0, 0.0);
                         for (size_t j = 0; j < (size_t(1) << kk); ++j) {
                              size_t loc = i;
                              for (size_t l = 0; l < kk; ++l)
                                   if (((j >> l) & 1) != ((i >> idx2mat[l]) & 1))
                                        loc ^= size_t(1) << idx2mat[l];
                              res += cmul(oldcol[loc], item.mat[li][j]);
                         }
                         M[i][col] = res;
                    }
               }
          }
          ctrls.assign(ctrl_set_.begin(), ctrl_set_.end());
          factor = 1.0;
     }
};

// ---------------------------------------------------------------------------
// Simulator: SimulatorMPI.{hpp,cpp} at world size 1.
// ---------------------------------------------------------------------------
struct Sim {
     size_t max_local, max_cluster, min_local;
     std::vector<Index> locals;
     std::vector<Complex> vec;
     Fusion fusion;
     std::mt19937 eng;
     std::uniform_real_distribution<double> dist{0.0, 1.0};

     // SimulatorMPI.cpp:66-104: state starts as [1]; mt19937(seed).
     Sim(uint64_t seed, size_t ml, size_t mc)
         : max_local(ml), max_cluster(mc), min_local(mc), vec(1, Complex(1.0, 0.0)),
           eng(static_cast<std::mt19937::result_type>(seed))
     {}

     size_t find(Index id) const
     {
          auto it = std::find(locals.begin(), locals.end(), id);
          return it == locals.end() ? kNotFound : static_cast<size_t>(it - locals.begin());
     }
     size_t find_sure(Index id) const
     {
          size_t p = find(id);
          if (p == kNotFound)
               throw std::runtime_error("ArrayFindSure(): Can't find " + std::to_string(id));
          return p;
     }

     // SimulatorMPI.cpp:185-216 (world size 1: no global slots).
     void allocate_qubit(Index id)
     {
          if (locals.size() < min_local || locals.size() < max_local) {
               locals.push_back(id);
               vec.resize(vec.size() * 2, Complex(0.0, 0.0));
          }
          else {
               throw std::runtime_error("AllocateQubit(): can't allocate more than "
                                        + std::to_string(locals.size()) + " qubits");
          }
     }

     // SimulatorMPI.cpp:218-251.
     void allocate_qureg(const std::vector<Index>& ids, Complex init)
     {
          if (init != 0.0 && !locals.empty())
               throw std::runtime_error(
                   "AllocateQureg(): initialization of only first qureg is supported");
          for (auto i: ids) allocate_qubit(i);
          if (init != 0.0) std::fill(vec.begin(), vec.end(), init);
     }

     // SimulatorMPI.cpp:278-322 (DeallocateLocalQubit) via :369-426.
     void deallocate_qubit(Index id)
     {
          size_t pos = find(id);
          if (pos == kNotFound)
               throw std::runtime_error("ArrayFindSure(): Can't find " + std::to_string(id));
          double sum[2] = {0.0, 0.0};
          for (size_t i = 0; i < vec.size(); ++i) sum[(i >> pos) & 1] += cnorm(vec[i]);
          if (!((sum[0] > kMaxFloatError) ^ (sum[1] > kMaxFloatError)))
               throw std::runtime_error("DeallocateLocalQubit(): qubit "
                                        + std::to_string(id) + " is entangled");
          const size_t blk = size_t(1) << pos;
          std::vector<Complex> nv(vec.size() / 2);
          const size_t keep = sum[0] > kMaxFloatError ? 0 : 1;
          for (size_t b = 0; b < nv.size() / blk; ++b)
               for (size_t j = 0; j < blk; ++j) nv[b * blk + j] = vec[(2 * b + keep) * blk + j];
          vec.swap(nv);
          locals.erase(locals.begin() + static_cast<long>(pos));
     }

     // SimulatorMPI.cpp:710-815 (world size 1: no rank pruning).
     void apply_gate(Matrix m, std::vector<Index> ids, const std::vector<Index>& ctrls)
     {
          if (ids.size() > 62 || m.size() != (size_t(1) << ids.size()))
               throw std::runtime_error("apply_controlled_gate(): matrix size does not match");
          bool diag = is_diag(m);
          std::vector<Index> lc;
          for (auto c: ctrls)
               if (find(c) != kNotFound) lc.push_back(c);
          if (ids.size() + lc.size() > max_cluster) {
               run();
               fusion.insert(m, diag, ids, ctrls);
               return;
          }
          Fusion::add_controls(m, ids, lc);
          fusion.insert(m, diag, ids, {});
     }

     // SimulatorMPI.cpp:441-541.
     void run()
     {
          Matrix m;
          std::vector<Index> ids, ctrls;
          bool diag = false;
          fusion.perform(m, ids, ctrls, diag);
          uint64_t cmask = 0;
          for (auto c: ctrls) {
               size_t p = find(c);
               if (p != kNotFound) cmask |= 1ull << p;
          }
          std::vector<uint32_t> pos(ids.size());
          for (size_t i = 0; i < ids.size(); ++i) pos[i] = static_cast<uint32_t>(find_sure(ids[i]));
          if (ids.empty()) {
               if (m[0][0] != Complex(1.0, 0.0)) scale_pass(vec.data(), vec.size(), m[0][0]);
          }
          else if (ids.size() <= 5) {
               if (diag) diag_pass(vec.data(), vec.size(), m, pos, cmask);
               else dense_pass(vec.data(), vec.size(), m, pos, cmask);
          }
          else {
               throw std::runtime_error("Run(): cannot apply " + std::to_string(ids.size())
                                        + " qubits gate");
          }
          fusion = Fusion();
     }

     // SimulatorMPI.cpp:841-870.
     double prob_internal(uint64_t msk, uint64_t val) const
     {
          double p = 0.0;
          for (size_t i = 0; i < vec.size(); ++i)
               if ((i & msk) == val) p += cnorm(vec[i]);
          return p;
     }

     // SimulatorMPI.cpp:872-895.
     void normalize(double norm, uint64_t msk, uint64_t val)
     {
          double f = 1.0 / std::sqrt(norm);
          for (size_t i = 0; i < vec.size(); ++i) {
               if ((i & msk) != val) vec[i] = 0.0;
               else vec[i] = Complex(vec[i].real() * f, vec[i].imag() * f);
          }
     }

     void mask_of(const std::vector<Index>& ids, const std::vector<uint8_t>& bits,
                  uint64_t& msk, uint64_t& val) const
     {
          msk = val = 0;
          for (size_t i = 0; i < ids.size(); ++i) {
               size_t p = find_sure(ids[i]);
               msk |= 1ull << p;
               if (bits[i]) val |= 1ull << p;
          }
     }

     // SimulatorMPI.cpp:817-839 + 897-1008.  One u per call; nb = min(size,
     // 2^15) contiguous blocks summed sequentially; partial_sum; first block
     // with u <= cum[b]; in-block scan from cum[b-1] until acc >= u (k may land
     // one past the block end if rounding keeps acc below u).
     std::vector<uint8_t> measure(const std::vector<Index>& ids)
     {
          const uint64_t nb = std::min<uint64_t>(vec.size(), uint64_t(1) << 15);
          const uint64_t K = vec.size() / nb;
          std::vector<double> cum(nb, 0.0);
          for (uint64_t b = 0; b < nb; ++b)
               for (uint64_t j = 0; j < K; ++j) cum[b] += cnorm(vec[b * K + j]);
          std::partial_sum(cum.begin(), cum.end(), cum.begin());
          double u = dist(eng);
          uint64_t b = 0;
          for (; b < nb; ++b)
               if (u <= cum[b]) break;
          if (b == nb)
               throw std::runtime_error("MeasureQubits(): Random number > norm partial sum");
          double shift = b > 0 ? cum[b - 1] : 0.0;
          uint64_t k = b * K;
          for (uint64_t j = 0; j < K; ++j, ++k) {
               shift += cnorm(vec[k]);
               if (shift >= u) break;
          }
          std::vector<uint8_t> res(ids.size(), 0);
          uint64_t msk = 0, val = 0;
          for (size_t i = 0; i < ids.size(); ++i) {
               size_t p = find_sure(ids[i]);
               msk |= 1ull << p;
               if (k & (1ull << p)) {
                    val |= 1ull << p;
                    res[i] = 1;
               }
          }
          double norm = prob_internal(msk, val);
          normalize(norm, msk, val);
          return res;
     }

     // SimulatorMPI.cpp:561-600.
     double get_probability(const std::vector<uint8_t>& bits, const std::vector<Index>& ids) const
     {
          if (ids.size() != bits.size())
               throw std::runtime_error("GetProbability(): ids.size() != bit_string.size()");
          uint64_t msk, val;
          mask_of(ids, bits, msk, val);
          return prob_internal(msk, val);
     }

     // SimulatorMPI.cpp:602-650.
     Complex get_amplitude(const std::vector<uint8_t>& bits, const std::vector<Index>& ids) const
     {
          if (bits.size() != locals.size() || bits.size() != ids.size())
               throw std::runtime_error("GetAmplitude(): ids.size() != number of qubits");
          uint64_t num = 0, check = 0;
          for (size_t i = 0; i < ids.size(); ++i) {
               size_t p = find_sure(ids[i]);
               if (bits[i]) num |= 1ull << p;
               check |= 1ull << p;
          }
          if (check != (uint64_t(1) << locals.size()) - 1)
               throw std::runtime_error(
                   "GetAmplitude(): the second argument must be a permutation of all "
                   "allocated qubits.");
          return vec[num];
     }

     // SimulatorMPI.cpp:1010-1058.
     void collapse(const std::vector<Index>& ids, const std::vector<uint8_t>& values)
     {
          if (ids.size() != values.size())
               throw std::runtime_error("collapseWaveFunction(): ids.size() != values.size()");
          uint64_t msk, val;
          mask_of(ids, values, msk, val);
          double norm = prob_internal(msk, val);
          if (norm < 1.e-12)
               throw std::runtime_error(
                   "collapseWaveFunction(): Invalid collapse! Probability is ~0.");
          normalize(norm, msk, val);
     }

     // SimulatorMPI.cpp:681-694.
     double entropy() const
     {
          double e = 0.0;
          for (auto& a: vec) {
               double pr = std::abs(a) * std::abs(a);
               if (pr > 0) e += pr * std::log2(pr);
          }
          return -e;
     }

     // New (no reference counterpart; the reference's qswap moves bits
     // between ranks, SimulatorMPI.cpp:1085-1138): move the data so that
     // new_order[p] is the qubit at bit p.  Closed form: out[sigma(i)] = in[i].
     void permute(const std::vector<Index>& new_order)
     {
          if (new_order.size() != locals.size())
               throw std::runtime_error("permute_qubits(): wrong number of qubits");
          std::vector<uint32_t> src(new_order.size());  // new bit p <- old bit src[p]
          uint64_t seen = 0;
          for (size_t p = 0; p < new_order.size(); ++p) {
               src[p] = static_cast<uint32_t>(find_sure(new_order[p]));
               seen |= 1ull << src[p];
          }
          if (seen != (uint64_t(1) << locals.size()) - 1)
               throw std::runtime_error("permute_qubits(): ids must be a permutation");
          std::vector<Complex> nv(vec.size());
          for (uint64_t j = 0; j < vec.size(); ++j) {
               uint64_t i = 0;
               for (size_t p = 0; p < src.size(); ++p) i |= ((j >> p) & 1ull) << src[p];
               nv[j] = vec[i];
          }
          vec.swap(nv);
          locals = new_order;
     }

     // New (declared but unbound in the reference, _simulator_mpi.py:148-183):
     // <psi| P |psi> for a Pauli string.  P|i> = phase(i)|i ^ x>, so
     // <psi|P|psi> = sum_i conj(psi[i ^ x]) * phase(i) * psi[i], where
     // X|b> = |1-b>, Y|b> = i(-1)^b |1-b>, Z|b> = (-1)^b |b>.
     Complex expectation_term(uint64_t xmask, uint64_t zmask) const
     {
          // Y on a bit = X and Z both set; Y = i X Z.
          const uint64_t ymask = xmask & zmask;
          const int ny = __builtin_popcountll(ymask);
          const Complex iy[4] = {Complex(1, 0), Complex(0, 1), Complex(-1, 0), Complex(0, -1)};
          Complex acc(0.0, 0.0);
          for (uint64_t i = 0; i < vec.size(); ++i) {
               // Applying Z-part first then X-part: P = i^ny X^x Z^z.
               double sgn = (__builtin_popcountll(i & zmask) & 1) ? -1.0 : 1.0;
               Complex ph = iy[ny & 3] * sgn;
               acc += cmul(std::conj(vec[i ^ xmask]), cmul(ph, vec[i]));
          }
          return acc;
     }

     // New (_simulator_mpi.py:185-223, unbound in the reference): psi <- sum_t c_t P_t psi.
     void apply_pauli_sum(const std::vector<uint64_t>& xm, const std::vector<uint64_t>& zm,
                          const std::vector<Complex>& coeff)
     {
          std::vector<Complex> out(vec.size(), Complex(0.0, 0.0));
          const Complex iy[4] = {Complex(1, 0), Complex(0, 1), Complex(-1, 0), Complex(0, -1)};
          for (size_t t = 0; t < xm.size(); ++t) {
               const int ny = __builtin_popcountll(xm[t] & zm[t]);
               for (uint64_t i = 0; i < vec.size(); ++i) {
                    double sgn = (__builtin_popcountll(i & zm[t]) & 1) ? -1.0 : 1.0;
                    Complex ph = iy[ny & 3] * sgn;
                    out[i ^ xm[t]] += cmul(coeff[t], cmul(ph, vec[i]));
               }
          }
          vec.swap(out);
     }
};

template <class F>
int guard(F&& f)
{
     try {
          f();
          return 0;
     } catch (const std::exception& e) {
          g_last_error = e.what();
          return 1;
     }
}

Matrix to_matrix(const double* m, size_t dim)
{
     Matrix M(dim, Row(dim));
     for (size_t i = 0; i < dim; ++i)
          for (size_t j = 0; j < dim; ++j)
               M[i][j] = Complex(m[2 * (i * dim + j)], m[2 * (i * dim + j) + 1]);
     return M;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_last_error.c_str(); }

// Kernel-level passes on a caller-owned interleaved complex128 array.
void orc_dense_pass(double* psi, uint64_t n, const double* m, int k, const uint32_t* pos,
                    uint64_t ctrlmask)
{
     dense_pass(reinterpret_cast<Complex*>(psi), n, to_matrix(m, size_t(1) << k),
                std::vector<uint32_t>(pos, pos + k), ctrlmask);
}
void orc_diag_pass(double* psi, uint64_t n, const double* m, int k, const uint32_t* pos,
                   uint64_t cmask)
{
     diag_pass(reinterpret_cast<Complex*>(psi), n, to_matrix(m, size_t(1) << k),
               std::vector<uint32_t>(pos, pos + k), cmask);
}

// Fusion-level: feed gates, then fuse.  Returns the fused width (or -1 if
// out_cap is too small), the fused matrix, ascending ids, ctrls, diag flag.
void* orc_fusion_new() { return new Fusion(); }
void orc_fusion_free(void* f) { delete static_cast<Fusion*>(f); }
void orc_fusion_insert(void* f, const double* m, size_t dim, const int64_t* ids, size_t nids,
                       const int64_t* ctrls, size_t nctrls, int diag)
{
     static_cast<Fusion*>(f)->insert(to_matrix(m, dim), diag != 0,
                                     std::vector<Index>(ids, ids + nids),
                                     std::vector<Index>(ctrls, ctrls + nctrls));
}
int orc_fusion_perform(void* f, double* out_m, size_t out_cap_dim, int64_t* out_ids,
                       int64_t* out_ctrls, size_t* nctrls, int* diag)
{
     Matrix M;
     std::vector<Index> ids, ctrls;
     bool d;
     static_cast<Fusion*>(f)->perform(M, ids, ctrls, d);
     if (M.size() > out_cap_dim) return -1;
     for (size_t i = 0; i < M.size(); ++i)
          for (size_t j = 0; j < M.size(); ++j) {
               out_m[2 * (i * M.size() + j)] = M[i][j].real();
               out_m[2 * (i * M.size() + j) + 1] = M[i][j].imag();
          }
     std::copy(ids.begin(), ids.end(), out_ids);
     std::copy(ctrls.begin(), ctrls.end(), out_ctrls);
     *nctrls = ctrls.size();
     *diag = d ? 1 : 0;
     return static_cast<int>(ids.size());
}

// First `count` draws of uniform_real_distribution<double>(0,1) over
// mt19937(seed): the measurement u stream (SimulatorMPI.cpp:88,100-101).
void orc_uniform_stream(uint64_t seed, double* out, size_t count)
{
     std::mt19937 e(static_cast<std::mt19937::result_type>(seed));
     std::uniform_real_distribution<double> d(0.0, 1.0);
     for (size_t i = 0; i < count; ++i) out[i] = d(e);
}
uint32_t orc_mt19937_nth(uint64_t seed, size_t nth)
{
     std::mt19937 e(static_cast<std::mt19937::result_type>(seed));
     uint32_t v = 0;
     for (size_t i = 0; i < nth; ++i) v = static_cast<uint32_t>(e());
     return v;
}

// Simulator-level API (mirrors the _cppsim_mpi method set).
void* orc_sim_new(uint64_t seed, int max_local, int max_cluster)
{
     return new Sim(seed, static_cast<size_t>(max_local), static_cast<size_t>(max_cluster));
}
void orc_sim_free(void* s) { delete static_cast<Sim*>(s); }
int orc_allocate_qubit(void* s, int64_t id)
{
     return guard([&] { static_cast<Sim*>(s)->allocate_qubit(id); });
}
int orc_allocate_qureg(void* s, const int64_t* ids, size_t n, double re, double im)
{
     return guard([&] {
          static_cast<Sim*>(s)->allocate_qureg(std::vector<Index>(ids, ids + n), Complex(re, im));
     });
}
int orc_deallocate_qubit(void* s, int64_t id)
{
     return guard([&] { static_cast<Sim*>(s)->deallocate_qubit(id); });
}
int orc_apply_controlled_gate(void* s, const double* m, size_t dim, const int64_t* ids,
                              size_t nids, const int64_t* ctrls, size_t nctrls)
{
     return guard([&] {
          static_cast<Sim*>(s)->apply_gate(to_matrix(m, dim), std::vector<Index>(ids, ids + nids),
                                           std::vector<Index>(ctrls, ctrls + nctrls));
     });
}
int orc_run(void* s)
{
     return guard([&] { static_cast<Sim*>(s)->run(); });
}
int orc_measure_qubits(void* s, const int64_t* ids, size_t n, uint8_t* out)
{
     return guard([&] {
          auto r = static_cast<Sim*>(s)->measure(std::vector<Index>(ids, ids + n));
          std::copy(r.begin(), r.end(), out);
     });
}
int orc_get_probability(void* s, const uint8_t* bits, const int64_t* ids, size_t n, double* out)
{
     return guard([&] {
          *out = static_cast<Sim*>(s)->get_probability(std::vector<uint8_t>(bits, bits + n),
                                                       std::vector<Index>(ids, ids + n));
     });
}
int orc_get_amplitude(void* s, const uint8_t* bits, size_t nbits, const int64_t* ids, size_t nids,
                      double* out)
{
     return guard([&] {
          Complex a = static_cast<Sim*>(s)->get_amplitude(
              std::vector<uint8_t>(bits, bits + nbits), std::vector<Index>(ids, ids + nids));
          out[0] = a.real();
          out[1] = a.imag();
     });
}
int orc_collapse_wavefunction(void* s, const int64_t* ids, const uint8_t* vals, size_t n)
{
     return guard([&] {
          static_cast<Sim*>(s)->collapse(std::vector<Index>(ids, ids + n),
                                         std::vector<uint8_t>(vals, vals + n));
     });
}
int orc_entropy(void* s, double* out)
{
     return guard([&] { *out = static_cast<Sim*>(s)->entropy(); });
}
size_t orc_num_qubits(void* s) { return static_cast<Sim*>(s)->locals.size(); }
void orc_get_qubit_ids(void* s, int64_t* out)
{
     auto& l = static_cast<Sim*>(s)->locals;
     std::copy(l.begin(), l.end(), out);
}
int orc_set_qubits_perm(void* s, const int64_t* ids, size_t n)
{
     return guard([&] {
          auto* sim = static_cast<Sim*>(s);
          if (n != sim->locals.size())
               throw std::runtime_error("set_qubits_perm(): wrong number of ids");
          sim->locals.assign(ids, ids + n);
     });
}
int orc_permute_qubits(void* s, const int64_t* ids, size_t n)
{
     return guard([&] { static_cast<Sim*>(s)->permute(std::vector<Index>(ids, ids + n)); });
}
uint64_t orc_state_size(void* s) { return static_cast<Sim*>(s)->vec.size(); }
void orc_get_state(void* s, double* out)
{
     auto& v = static_cast<Sim*>(s)->vec;
     std::memcpy(out, v.data(), v.size() * sizeof(Complex));
}
int orc_set_state(void* s, const double* in, uint64_t n)
{
     return guard([&] {
          auto& v = static_cast<Sim*>(s)->vec;
          if (n != v.size()) throw std::runtime_error("set_state(): size mismatch");
          std::memcpy(static_cast<void*>(v.data()), in, n * sizeof(Complex));
     });
}
int orc_expectation_term(void* s, uint64_t xmask, uint64_t zmask, double* out)
{
     return guard([&] {
          Complex r = static_cast<Sim*>(s)->expectation_term(xmask, zmask);
          out[0] = r.real();
          out[1] = r.imag();
     });
}
int orc_apply_pauli_sum(void* s, const uint64_t* xm, const uint64_t* zm, const double* coeff,
                        size_t nterms)
{
     return guard([&] {
          std::vector<Complex> c(nterms);
          for (size_t t = 0; t < nterms; ++t) c[t] = Complex(coeff[2 * t], coeff[2 * t + 1]);
          static_cast<Sim*>(s)->apply_pauli_sum(std::vector<uint64_t>(xm, xm + nterms),
                                                std::vector<uint64_t>(zm, zm + nterms), c);
     });
}

}  // extern "C"
