// Gate schedulers; see sched.hpp for the contract.
#include "sched.hpp"

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>

namespace hq {
namespace {

inline int popc(msk_t x) { return __builtin_popcountll(x); }
inline bool submask(msk_t a, msk_t b) { return (a & ~b) == 0; }

// Dense renumbering of every id that appears (convertors.cpp:42-76).
struct Positions {
     std::vector<int64_t> pos_to_id;
     std::map<int64_t, int> id_to_pos;

     Positions(const GateList& g, const std::vector<int64_t>& a, const std::vector<int64_t>& b)
     {
          for (auto& v: g.gate) pos_to_id.insert(pos_to_id.end(), v.begin(), v.end());
          for (auto& v: g.ctrl) pos_to_id.insert(pos_to_id.end(), v.begin(), v.end());
          pos_to_id.insert(pos_to_id.end(), a.begin(), a.end());
          pos_to_id.insert(pos_to_id.end(), b.begin(), b.end());
          std::sort(pos_to_id.begin(), pos_to_id.end());
          pos_to_id.erase(std::unique(pos_to_id.begin(), pos_to_id.end()), pos_to_id.end());
          if (pos_to_id.size() >= 64)
               throw std::runtime_error(
                   "CalcPos(): scheduling with 64 or more qubits is not yet supported");
          for (size_t i = 0; i < pos_to_id.size(); ++i) id_to_pos[pos_to_id[i]] = static_cast<int>(i);
     }
     msk_t mask(const std::vector<int64_t>& ids) const
     {
          msk_t m = 0;
          for (auto id: ids) m |= msk_t(1) << id_to_pos.at(id);
          return m;
     }
};

void check_sizes(const GateList& g)
{
     if (g.gate.size() != g.ctrl.size() || g.gate.size() != g.diag.size())
          throw std::runtime_error("scheduler: gate, gate_ctrl and gate_diag differ in length");
}

// Open-addressing map msk_t -> msk_t (the "used" bits per cluster).  Key 0
// is a valid cluster, so occupancy is tracked separately.
class MaskMap {
public:
     MaskMap() { rehash(1 << 12); }
     msk_t& at(msk_t key)
     {
          if ((size_ + 1) * 2 > keys_.size()) rehash(keys_.size() * 2);
          size_t i = slot(key);
          if (!full_[i]) {
               full_[i] = 1;
               keys_[i] = key;
               vals_[i] = 0;
               ++size_;
          }
          return vals_[i];
     }

private:
     size_t slot(msk_t key) const
     {
          const size_t mask = keys_.size() - 1;
          size_t i = static_cast<size_t>((key * 0x9E3779B97F4A7C15ull) >> 20) & mask;
          while (full_[i] && keys_[i] != key) i = (i + 1) & mask;
          return i;
     }
     void rehash(size_t cap)
     {
          std::vector<msk_t> ok = std::move(keys_), ov = std::move(vals_);
          std::vector<uint8_t> of = std::move(full_);
          keys_.assign(cap, 0);
          vals_.assign(cap, 0);
          full_.assign(cap, 0);
          size_ = 0;
          for (size_t i = 0; i < of.size(); ++i)
               if (of[i]) {
                    size_t j = slot(ok[i]);
                    full_[j] = 1;
                    keys_[j] = ok[i];
                    vals_[j] = ov[i];
                    ++size_;
               }
     }
     std::vector<msk_t> keys_, vals_;
     std::vector<uint8_t> full_;
     size_t size_ = 0;
};

class Cluster {
public:
     Cluster(const GateList& g, const std::vector<int64_t>& locals,
             const std::vector<int64_t>& globals, int cluster_size)
         : size_(cluster_size), diag_(g.diag)
     {
          check_sizes(g);
          Positions p(g, locals, globals);
          for (size_t i = 0; i < g.gate.size(); ++i) {
               gate_.push_back(p.mask(g.gate[i]));
               ctrl_.push_back(p.mask(g.ctrl[i]));
          }
          locals_ = p.mask(locals);
          // nonlocal_after_[i]: some gate at index >= i has no local qubit.
          nonlocal_after_.assign(gate_.size() + 1, 0);
          for (size_t i = gate_.size(); i-- > 0;)
               nonlocal_after_[i] = nonlocal_after_[i + 1] || (((gate_[i] | ctrl_[i]) & locals_) == 0);
     }

     std::vector<int> run()
     {
          if (gate_.empty()) return {};
          best_ans_ = 0;
          best_ = 0;
          for (size_t i = 0; i < gate_.size(); ++i) {
               const msk_t cur = locals_ & (gate_[i] | ctrl_[i]);
               if (popc(cur) <= size_) rec(cur, 1);
          }
          if (best_ans_ != 0) return take(best_);
          msk_t bad = 0;
          for (size_t i = 0; i < gate_.size(); ++i) {
               if (can_take(i, locals_, bad)) return {static_cast<int>(i)};
               bad |= gate_[i] | ctrl_[i];
          }
          return {};
     }

private:
     bool can_take(size_t i, msk_t cur, msk_t bad) const
     {
          const msk_t all = gate_[i] | ctrl_[i];
          if ((all & bad) || !submask(all & locals_, cur)) return false;
          return diag_[i] ? true : submask(gate_[i], locals_);
     }

     // cluster_scheduler.h:76-99 (CalcCombination).
     void evaluate(msk_t cur)
     {
          msk_t bad = 0;
          int ans = 0;
          for (size_t i = 0; i < gate_.size(); ++i) {
               if (can_take(i, cur, bad)) ++ans;
               else bad |= gate_[i] | ctrl_[i];
               if (submask(cur, bad) && !nonlocal_after_[i + 1]) break;
          }
          if (ans > best_ans_ || (ans == best_ans_ && popc(cur) < popc(best_))) {
               best_ans_ = ans;
               best_ = cur;
          }
     }

     // cluster_scheduler.cpp:80-99 (Rec); the memo records, per cluster,
     // which "next bit" expansions were done; a cluster is evaluated once.
     void rec(msk_t cur, msk_t bit)
     {
          msk_t& used = memo_.at(cur);
          if (used == 0) evaluate(cur);
          if (bit > locals_) {
               if (used == 0) used = msk_t(1) << 63;  // evaluated, nothing expanded
               return;
          }
          if (used & bit) return;
          used |= bit;
          if (submask(bit, locals_) && !(bit & cur) && popc(cur) < size_) rec(cur | bit, bit << 1);
          rec(cur, bit << 1);
     }

     std::vector<int> take(msk_t cluster) const
     {
          std::vector<int> ans;
          msk_t bad = 0;
          for (size_t i = 0; i < gate_.size(); ++i) {
               if (can_take(i, cluster, bad)) ans.push_back(static_cast<int>(i));
               else bad |= gate_[i] | ctrl_[i];
          }
          return ans;
     }

     int size_;
     std::vector<bool> diag_;
     std::vector<msk_t> gate_, ctrl_;
     std::vector<uint8_t> nonlocal_after_;
     msk_t locals_ = 0;
     int best_ans_ = 0;
     msk_t best_ = 0;
     MaskMap memo_;
};

class Swap {
public:
     Swap(const GateList& g, int64_t num_splits, int num_locals, bool fuse)
         : num_splits_(num_splits), num_locals_(num_locals), diag_(g.diag), weight_(g.gate.size(), 1),
           pos_(g, {}, {})
     {
          check_sizes(g);
          for (size_t i = 0; i < g.gate.size(); ++i) {
               gate_.push_back(pos_.mask(g.gate[i]));
               ctrl_.push_back(pos_.mask(g.ctrl[i]));
          }
          if (fuse) fuse_single_qubit_gates();
     }

     std::vector<int64_t> run()
     {
          if (gate_.empty()) return {};
          best_ans_ = 0;
          best_ = 0;
          rec(0, 0, 0, 0, num_splits_);
          std::vector<int64_t> ids;
          for (int i = 0; i < 64 && (msk_t(1) << i) <= best_; ++i)
               if ((best_ >> i) & 1) ids.push_back(pos_.pos_to_id[i]);
          return ids;
     }

private:
     bool can_take(size_t i, msk_t cur, msk_t bad) const
     {
          if ((ctrl_[i] | gate_[i]) & bad) return false;
          return diag_[i] ? true : popc(gate_[i] | cur) <= num_locals_;
     }

     // swap_scheduler.h:87-104 (FuseSingleQubitGateToThis).
     void fuse_into(size_t i, size_t j)
     {
          weight_[j] += weight_[i];
          weight_[i] = 0;
          if (!diag_[i]) diag_[j] = false;
          if (ctrl_[j] & gate_[i]) {
               ctrl_[j] ^= gate_[i];
               gate_[j] |= gate_[i];
          }
     }
     bool fuse_prev(size_t i)
     {
          for (size_t j = i; j-- > 0;)
               if (gate_[i] & (gate_[j] | ctrl_[j])) {
                    fuse_into(i, j);
                    return true;
               }
          return false;
     }
     bool fuse_next(size_t i)
     {
          for (size_t j = i + 1; j < gate_.size(); ++j)
               if (gate_[i] & (gate_[j] | ctrl_[j])) {
                    fuse_into(i, j);
                    return true;
               }
          return false;
     }
     // swap_scheduler.cpp:47-68.
     void fuse_single_qubit_gates()
     {
          for (size_t i = gate_.size(); i-- > 0;) {
               bool remove = false;
               if (popc(gate_[i] | ctrl_[i]) == 1)
                    remove = !diag_[i] ? (fuse_next(i) || fuse_prev(i)) : (fuse_prev(i) || fuse_next(i));
               if (remove) {
                    gate_.erase(gate_.begin() + static_cast<long>(i));
                    ctrl_.erase(ctrl_.begin() + static_cast<long>(i));
                    diag_.erase(diag_.begin() + static_cast<long>(i));
                    weight_.erase(weight_.begin() + static_cast<long>(i));
               }
          }
     }

     // swap_scheduler.cpp:90-170: take/skip backtracking with a split budget.
     int64_t rec(size_t pos, msk_t cur, msk_t bad, int64_t ans, int64_t splits)
     {
          if (ans > best_ans_) {
               best_ans_ = ans;
               best_ = cur;
          }
          if (pos == gate_.size()) return splits;
          bool can_skip = true;
          int cnt = 1;
          const bool can_take_now = can_take(pos, cur, bad);
          if (can_take_now) {
               if (!diag_[pos]) {
                    if (submask(gate_[pos], cur)) {
                         can_skip = false;
                         --cnt;
                    }
                    ++cnt;
               }
               else {
                    can_skip = false;
               }
          }
          if (!(0 < cnt && cnt <= 2)) throw std::logic_error("SwapScheduler: branch count");
          if (splits == 0 && can_take_now && can_skip) {
               can_skip = false;
               --cnt;
          }
          splits -= cnt - 1;
          if (splits < 0) throw std::logic_error("SwapScheduler: split budget");
          if (can_skip) {
               const int64_t give = splits / cnt;
               splits += rec(pos + 1, cur, bad | gate_[pos] | ctrl_[pos], ans, give) - give;
               --cnt;
          }
          if (can_take_now) {
               const msk_t nl = diag_[pos] ? cur : (cur | gate_[pos]);
               splits = rec(pos + 1, nl, bad, ans + weight_[pos], splits);
               --cnt;
          }
          if (cnt != 0) throw std::logic_error("SwapScheduler: invalid branch number");
          return splits;
     }

     int64_t num_splits_;
     int num_locals_;
     std::vector<bool> diag_;
     std::vector<int64_t> weight_;
     Positions pos_;
     std::vector<msk_t> gate_, ctrl_;
     int64_t best_ans_ = 0;
     msk_t best_ = 0;
};

}  // namespace

std::vector<int> schedule_cluster(const GateList& g, const std::vector<int64_t>& locals,
                                  const std::vector<int64_t>& globals, int cluster_size)
{
     Cluster c(g, locals, globals, cluster_size);
     return c.run();
}

std::vector<int> plan_clusters(const GateList& g, int cluster_size, int mode)
{
     check_sizes(g);
     if (mode != kPlanReference && mode != kPlanLookahead)
          throw std::runtime_error("plan_clusters(): mode must be 0 (reference) or 1 (lookahead)");
     if (cluster_size < 1) throw std::runtime_error("plan_clusters(): cluster_size must be >= 1");
     const size_t ng = g.gate.size();
     std::vector<int> cluster_of(ng, -1);
     if (ng == 0) return cluster_of;
     if (mode == kPlanReference) {
          std::vector<int64_t> all;
          for (size_t i = 0; i < ng; ++i) {
               all.insert(all.end(), g.gate[i].begin(), g.gate[i].end());
               all.insert(all.end(), g.ctrl[i].begin(), g.ctrl[i].end());
          }
          std::sort(all.begin(), all.end());
          all.erase(std::unique(all.begin(), all.end()), all.end());
          std::vector<size_t> remaining(ng);
          for (size_t i = 0; i < ng; ++i) remaining[i] = i;
          int cid = 0;
          while (!remaining.empty()) {
               GateList sub;
               for (auto i: remaining) {
                    sub.gate.push_back(g.gate[i]);
                    sub.ctrl.push_back(g.ctrl[i]);
                    sub.diag.push_back(false);
               }
               auto avail = schedule_cluster(sub, all, {}, cluster_size);
               if (avail.empty()) throw std::runtime_error("plan_clusters(): no cluster can be formed");
               std::vector<uint8_t> taken(remaining.size(), 0);
               for (int a: avail) {
                    taken[a] = 1;
                    cluster_of[remaining[a]] = cid;
               }
               std::vector<size_t> next;
               for (size_t j = 0; j < remaining.size(); ++j)
                    if (!taken[j]) next.push_back(remaining[j]);
               remaining.swap(next);
               ++cid;
          }
          return cluster_of;
     }
     // Lookahead packing on dense bit masks of the qubit ids.
     Positions pos(g, {}, {});
     std::vector<msk_t> q(ng);
     for (size_t i = 0; i < ng; ++i) q[i] = pos.mask(g.gate[i]) | pos.mask(g.ctrl[i]);
     std::vector<size_t> remaining(ng);
     for (size_t i = 0; i < ng; ++i) remaining[i] = i;
     // Gates a fixed qubit set S can take from the front of `remaining`
     // (a gate is takeable if it lies in S and touches no qubit of a gate
     // skipped before it); the scan ends once every qubit of S is blocked.
     auto score = [&](msk_t S) {
          msk_t blocked = 0;
          int n = 0;
          for (size_t j = 0; j < remaining.size(); ++j) {
               const msk_t m = q[remaining[j]];
               if (!(m & blocked) && submask(m, S)) ++n;
               else blocked |= m;
               if (submask(S, blocked)) break;
          }
          return n;
     };
     auto better = [](int n, msk_t S, int best_n, msk_t best) {
          return n > best_n || (n == best_n && popc(S) < popc(best));
     };
     const int kWindow = 8 * cluster_size + 16;  // leading gates whose qubits form the pool
     const int kSeeds = 16;
     auto pool_gates = [](size_t j, int window) { return j < static_cast<size_t>(window); };
     int cid = 0;
     while (!remaining.empty()) {
          msk_t best = 0;
          int best_n = -1;
          msk_t pool = 0;
          {
               // Seeds: the first gates that are still takeable (no earlier
               // gate shares a qubit).  Each seed grows one qubit at a time,
               // adding the pool qubit that lets the most gates be taken.
               msk_t blocked = 0;
               std::vector<msk_t> seeds;
               for (size_t j = 0; j < remaining.size() && pool_gates(j, kWindow); ++j) {
                    const msk_t m0 = q[remaining[j]];
                    pool |= m0;
                    if (!(m0 & blocked) && static_cast<int>(seeds.size()) < kSeeds)
                         seeds.push_back(m0);
                    blocked |= m0;
               }
               auto offer = [&](msk_t S) {
                    const int n = score(S);
                    if (better(n, S, best_n, best)) {
                         best_n = n;
                         best = S;
                    }
               };
               for (msk_t seed: seeds) {
                    // (a) first-come: take gates in program order while they fit
                    msk_t S = seed, bl = 0;
                    for (size_t i = 0; i < remaining.size() && popc(S) < cluster_size; ++i) {
                         const msk_t m = q[remaining[i]];
                         if (!(m & bl) && popc(S | m) <= cluster_size) S |= m;
                         else bl |= m;
                         if (submask(S, bl)) break;
                    }
                    offer(S);
                    // (b) best-add: grow by the pool qubit that takes most gates
                    S = seed;
                    while (popc(S) < cluster_size) {
                         msk_t add = 0;
                         int add_n = -1;
                         for (msk_t in = pool & ~S; in; in &= in - 1) {
                              const msk_t b = in & (~in + 1);
                              const int nn = score(S | b);
                              if (nn > add_n) {
                                   add_n = nn;
                                   add = b;
                              }
                         }
                         if (!add) break;
                         S |= add;
                    }
                    offer(S);
               }
          }
          // Hill climb: replace one qubit of the set by one of the pool.
          for (int round = 0; round < 8 && popc(best) <= cluster_size; ++round) {
               msk_t cand_best = best;
               int cand_n = best_n;
               for (msk_t out = best; out; out &= out - 1) {
                    const msk_t bo = out & (~out + 1);
                    for (msk_t in = pool & ~best; in; in &= in - 1) {
                         const msk_t S = (best & ~bo) | (in & (~in + 1));
                         const int n = score(S);
                         if (better(n, S, cand_n, cand_best)) {
                              cand_n = n;
                              cand_best = S;
                         }
                    }
                    if (popc(best) > 1) {
                         const msk_t S = best & ~bo;
                         const int n = score(S);
                         if (better(n, S, cand_n, cand_best)) {
                              cand_n = n;
                              cand_best = S;
                         }
                    }
               }
               if (popc(best) < cluster_size)
                    for (msk_t in = pool & ~best; in; in &= in - 1) {
                         const msk_t S = best | (in & (~in + 1));
                         const int n = score(S);
                         if (better(n, S, cand_n, cand_best)) {
                              cand_n = n;
                              cand_best = S;
                         }
                    }
               if (cand_best == best) break;
               best = cand_best;
               best_n = cand_n;
          }
          // Take the gates of the chosen set.
          msk_t blocked = 0;
          std::vector<size_t> next;
          next.reserve(remaining.size());
          for (size_t j = 0; j < remaining.size(); ++j) {
               const msk_t m = q[remaining[j]];
               if (!(m & blocked) && submask(m, best)) cluster_of[remaining[j]] = cid;
               else {
                    blocked |= m;
                    next.push_back(remaining[j]);
               }
          }
          if (next.size() == remaining.size())
               throw std::runtime_error("plan_clusters(): no gate could be scheduled");
          remaining.swap(next);
          ++cid;
     }
     return cluster_of;
}

TilePassPlanner::TilePassPlanner(const GateList& g, int n, int tile_high, int op_width)
    : g_(g), n_(n), tile_high_(tile_high), op_width_(op_width)
{
     check_sizes(g_);
     if (n < 6 + tile_high || n > 63) throw std::runtime_error("plan_tile_passes(): need 6 + tile_high <= n <= 63");
     if (tile_high < 1 || op_width < 1) throw std::runtime_error("plan_tile_passes(): bad tile or op width");
     const size_t ng = g_.gate.size();
     q_.resize(ng);
     for (size_t i = 0; i < ng; ++i) {
          msk_t m = 0;
          for (auto p: g_.gate[i]) m |= msk_t(1) << p;
          for (auto p: g_.ctrl[i]) m |= msk_t(1) << p;
          for (auto p: g_.gate[i])
               if (p < 0 || p >= n) throw std::runtime_error("plan_tile_passes(): position out of range");
          for (auto p: g_.ctrl[i])
               if (p < 0 || p >= n) throw std::runtime_error("plan_tile_passes(): position out of range");
          if (popc(m & ~msk_t(63)) > tile_high) throw std::runtime_error("plan_tile_passes(): a gate does not fit the tile");
          q_[i] = m;
     }
     remaining_.resize(ng);
     for (size_t i = 0; i < ng; ++i) remaining_[i] = i;
}

// Gates of the remaining list a tile over S would take.
int TilePassPlanner::score(uint64_t S) const
{
     msk_t blocked = 0;
     int cnt = 0;
     for (size_t j = 0; j < remaining_.size(); ++j) {
          const msk_t m = q_[remaining_[j]];
          if (!(m & blocked) && submask(m, S)) ++cnt;
          else blocked |= m;
          if (submask(S, blocked)) break;
     }
     return cnt;
}

bool TilePassPlanner::next(TilePlan& out)
{
     if (remaining_.empty()) return false;
     const msk_t kFree = 63;
     const size_t kWindow = 64;
     // pool: high positions of the leading gates, in order of first use
     std::vector<int> pool;
     msk_t seen = 0;
     for (size_t j = 0; j < remaining_.size() && j < kWindow; ++j)
          for (msk_t m = q_[remaining_[j]] & ~kFree & ~seen; m; m &= m - 1) {
               pool.push_back(__builtin_ctzll(m));
               seen |= m & (~m + 1);
          }
     msk_t H = q_[remaining_[0]] & ~kFree;  // the first gate always fits
     int best = score(kFree | H);
     while (popc(H) < tile_high_) {
          int add = -1, add_s = -1;
          for (int p: pool) {
               if ((H >> p) & 1) continue;
               const int s = score(kFree | H | (msk_t(1) << p));
               if (s > add_s) {
                    add_s = s;
                    add = p;
               }
          }
          if (add < 0) break;
          H |= msk_t(1) << add;
          best = add_s;
     }
     // one round of single swaps
     for (msk_t out_b = H; out_b; out_b &= out_b - 1) {
          const msk_t ob = out_b & (~out_b + 1);
          for (int p: pool) {
               if ((H >> p) & 1) continue;
               const msk_t H2 = (H & ~ob) | (msk_t(1) << p);
               const int s = score(kFree | H2);
               if (s > best) {
                    best = s;
                    H = H2;
                    break;
               }
          }
     }
     const msk_t S = kFree | H;
     // taken gates
     msk_t blocked = 0;
     std::vector<size_t> taken, next;
     for (size_t j = 0; j < remaining_.size(); ++j) {
          const msk_t m = q_[remaining_[j]];
          if (!(m & blocked) && submask(m, S)) taken.push_back(remaining_[j]);
          else {
               blocked |= m;
               next.push_back(remaining_[j]);
          }
     }
     if (taken.empty()) throw std::runtime_error("plan_tile_passes(): no gate could be scheduled");
     remaining_.swap(next);
     // pad the high positions to tile_high with the lowest free ones
     for (int p = 6; p < n_ && popc(H) < tile_high_; ++p) H |= msk_t(1) << p;
     out = TilePlan{};
     for (int p = 6; p < n_; ++p)
          if ((H >> p) & 1) out.hpos.push_back(p);
     // ops: lookahead packing of the taken gates
     GateList sub;
     for (size_t i: taken) {
          sub.gate.push_back(g_.gate[i]);
          sub.ctrl.push_back(g_.ctrl[i]);
          sub.diag.push_back(false);
     }
     const std::vector<int> cl = plan_clusters(sub, op_width_, kPlanLookahead);
     const int nc = *std::max_element(cl.begin(), cl.end()) + 1;
     out.ops.assign(nc, {});
     for (size_t j = 0; j < taken.size(); ++j) out.ops[cl[j]].push_back(static_cast<int>(taken[j]));
     return true;
}

std::vector<TilePlan> plan_tile_passes(const GateList& g, int n, int tile_high, int op_width)
{
     TilePassPlanner planner(g, n, tile_high, op_width);
     std::vector<TilePlan> out;
     TilePlan tp;
     while (planner.next(tp)) out.push_back(std::move(tp));
     return out;
}

std::vector<int64_t> schedule_swap(const GateList& g, int64_t num_splits, int num_locals, bool fuse)
{
     Swap s(g, num_splits, num_locals, fuse);
     return s.run();
}

}  // namespace hq
