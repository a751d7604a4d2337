// Gate fusion for the state-vector engine (host side, FP64).
//
// Semantics follow fusion_mpi.hpp:64-234 of the reference, because the fused
// matrix and its qubit order are part of the observable contract
// (test_simulator_kqubit_gate pins them):
//  * a 1x1 matrix multiplies a global factor (fusion_mpi.hpp:87-90);
//  * the fused qubit order is ascending qubit id, fused bit l <-> l-th
//    smallest id (fusion_mpi.hpp:104, std::set order);
//  * fused M = G_last ... G_1 (factor I) (fusion_mpi.hpp:120-158);
//  * the IS_DIAG flag survives only if every item is diagonal (:117-119);
//  * controls shared by the first item stay a control mask (ctrl_set_), a
//    control not shared by every item is folded into the items as the
//    highest matrix bits, diag(I, U) (fusion_mpi.hpp:167-226).
// The expansion of an item into the fused space is done here by applying the
// item's 2^kk x 2^kk matrix to every column of M as a kk-qubit gate on an
// N-qubit vector (one strided pass per column), rather than the reference's
// per-element index rebuild.
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <set>
#include <vector>

namespace hq {

using cplx = std::complex<double>;

// Dense square matrix, row-major.
struct Mat {
     size_t dim = 0;
     std::vector<cplx> a;
     Mat() = default;
     explicit Mat(size_t d) : dim(d), a(d * d, cplx(0.0, 0.0)) {}
     cplx& operator()(size_t i, size_t j) { return a[i * dim + j]; }
     const cplx& operator()(size_t i, size_t j) const { return a[i * dim + j]; }
};

// funcs.hpp:325-337: diagonal iff every off-diagonal entry is exactly 0.
inline bool is_diagonal(const Mat& m)
{
     for (size_t i = 0; i < m.dim; ++i)
          for (size_t j = 0; j < m.dim; ++j)
               if (i != j && m(i, j) != 0.0) return false;
     return true;
}

// diag(I_{F d - d}, U) with the new controls appended as the highest bits.
inline void add_controls(Mat& m, std::vector<int64_t>& ids, const std::vector<int64_t>& ctrls)
{
     if (ctrls.empty()) return;
     ids.insert(ids.end(), ctrls.begin(), ctrls.end());
     const size_t d = m.dim, nd = d << ctrls.size(), offset = nd - d;
     Mat r(nd);
     for (size_t i = 0; i < offset; ++i) r(i, i) = 1.0;
     for (size_t i = 0; i < d; ++i)
          for (size_t j = 0; j < d; ++j) r(offset + i, offset + j) = m(i, j);
     m = std::move(r);
}

class Fusion {
public:
     size_t num_qubits() const { return set_.size(); }
     size_t num_items() const { return items_.size(); }
     bool empty() const { return items_.empty() && factor_ == cplx(1.0, 0.0); }

     void insert(Mat m, bool diag, std::vector<int64_t> ids, const std::vector<int64_t>& ctrls)
     {
          if (m.dim == 1) {
               factor_ *= m(0, 0);
               return;
          }
          for (auto q: ids) set_.insert(q);
          handle_controls(m, ids, ctrls);
          items_.push_back(Item{std::move(m), std::move(ids), diag});
     }

     // Produces the fused matrix over ascending ids, the shared controls and
     // the diagonal flag; resets the factor (fusion_mpi.hpp:163).
     void perform(Mat& M, std::vector<int64_t>& ids, std::vector<int64_t>& ctrls, bool& diag)
     {
          ids.assign(set_.begin(), set_.end());
          const size_t N = ids.size();
          const size_t dim = size_t(1) << N;
          M = Mat(dim);
          for (size_t i = 0; i < dim; ++i) M(i, i) = factor_;
          diag = true;
          std::vector<cplx> col(dim), v;
          for (const auto& it: items_) {
               if (!it.diag) diag = false;
               const size_t kk = it.ids.size();
               std::vector<size_t> bit(kk);
               for (size_t l = 0; l < kk; ++l)
                    bit[l] = static_cast<size_t>(
                        std::lower_bound(ids.begin(), ids.end(), it.ids[l]) - ids.begin());
               const size_t gd = size_t(1) << kk;
               size_t gmask = 0;
               for (auto b: bit) gmask |= size_t(1) << b;
               std::vector<size_t> off(gd, 0);
               for (size_t j = 0; j < gd; ++j)
                    for (size_t l = 0; l < kk; ++l)
                         if ((j >> l) & 1) off[j] |= size_t(1) << bit[l];
               v.assign(gd, cplx(0.0, 0.0));
               for (size_t c = 0; c < dim; ++c) {
                    for (size_t r = 0; r < dim; ++r) col[r] = M(r, c);
                    for (size_t base = 0; base < dim; ++base) {
                         if (base & gmask) continue;
                         for (size_t j = 0; j < gd; ++j) v[j] = col[base | off[j]];
                         for (size_t i = 0; i < gd; ++i) {
                              cplx acc(0.0, 0.0);
                              for (size_t j = 0; j < gd; ++j) acc += it.m(i, j) * v[j];
                              M(base | off[i], c) = acc;
                         }
                    }
               }
          }
          ctrls.assign(ctrl_set_.begin(), ctrl_set_.end());
          factor_ = 1.0;
     }

     // Controls every item shares (kept as a control mask by perform()).
     std::vector<int64_t> shared_controls() const { return {ctrl_set_.begin(), ctrl_set_.end()}; }

     // Hands the pending items over unfused, in insertion order: each item's
     // matrix on its ids (controls not shared by every item already folded
     // in as diag(I, U)); the global factor is multiplied into the first
     // item.  The shared controls are those of shared_controls().  Leaves
     // the fusion empty.  Needs at least one item.
     void take_items(std::vector<Mat>& ms, std::vector<std::vector<int64_t>>& ids)
     {
          ms.clear();
          ids.clear();
          for (auto& it: items_) {
               ms.push_back(std::move(it.m));
               ids.push_back(std::move(it.ids));
          }
          if (!ms.empty() && factor_ != cplx(1.0, 0.0))
               for (auto& x: ms.front().a) x *= factor_;
          *this = Fusion();
     }

private:
     struct Item {
          Mat m;
          std::vector<int64_t> ids;
          bool diag;
     };

     void handle_controls(Mat& m, std::vector<int64_t>& ids, const std::vector<int64_t>& ctrls)
     {
          std::set<int64_t> unhandled = ctrl_set_;
          for (auto c: ctrls) {
               if (ctrl_set_.count(c)) {
                    unhandled.erase(c);
               }
               else if (!items_.empty()) {
                    add_controls(m, ids, {c});
                    set_.insert(c);
               }
               else {
                    ctrl_set_.insert(c);
               }
          }
          if (!unhandled.empty()) {
               std::vector<int64_t> nc(unhandled.begin(), unhandled.end());
               for (auto c: nc) {
                    ctrl_set_.erase(c);
                    set_.insert(c);
               }
               for (auto& it: items_) add_controls(it.m, it.ids, nc);
          }
     }

     std::set<int64_t> set_;
     std::vector<Item> items_;
     std::set<int64_t> ctrl_set_;
     cplx factor_ = 1.0;
};

}  // namespace hq
