// Many independent circuits in one device arena.
//
// The reference runs one circuit per simulator object: gates are queued
// (SimulatorMPI::ApplyGate), fused and applied on run() (SimulatorMPI::Run,
// SimulatorMPI.cpp:407-436).  A batch holds many such circuits, each from
// |0...0> on its own qubits 0..n-1 (bit p of a state's index = qubit p), and
// applies all of them in one graph launch:
//  * every circuit is cut into fused clusters (plan_clusters, lookahead
//    packer, at most k qubits per cluster with k chosen per qubit count) and
//    each cluster becomes one dense 2^k x 2^k pass (controls folded into the
//    matrix as diag(I, U) on the control bits);
//  * small states are owned by one workgroup each inside one persistent
//    launch: states of up to 128 KiB live in LDS for the whole circuit and
//    are written to HBM once; larger ones up to a size limit are streamed by
//    their workgroup pass after pass;
//  * large states are advanced layer by layer: launch L applies the L-th
//    pass of every large circuit (one launch per fused width), many
//    workgroups per state;
//  * the whole sequence is captured once into a hipGraph and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <vector>

#include "kernels.hpp"
#include "sched.hpp"

namespace hq {

// ---- device descriptors ---------------------------------------------------
// One fused pass of a workgroup-owned circuit.
struct BatchPass {
     uint64_t mat_off;  // first complex entry of the 2^k x 2^k matrix (row-major)
     uint8_t k;
     uint8_t pos[5];  // ascending bit positions; matrix bit l <-> pos[l]
     uint8_t pad[2];
};

// A workgroup-owned circuit.
struct BatchCirc {
     uint64_t state_off;  // first amplitude of the state in the arena
     uint32_t pass_begin, pass_end;
     uint32_t n;
     uint32_t pad;
};

// One pass of one large circuit inside a layered launch; items of a launch
// share k and lie back to back in blocks of kBatchLayerBlock groups.
struct BatchItem {
     uint64_t state_off;
     uint64_t mat_off;
     uint64_t block_start;  // first block of the item within the launch
     uint32_t log_groups;   // log2 of the 2^(n-k) groups
     uint8_t pos[5];        // ascending
     uint8_t k;             // fused width (launches that mix widths read it)
     uint8_t pad[2];
};

// One pass of one large circuit in the wave-tile form: each wave owns 64
// groups whose 64 * 2^k amplitudes are 2^k contiguous 64-amplitude chunks
// at base + lane + hoff[i]; they are regrouped through LDS (slot r of the
// tile = index bits at the tile positions, rank_n / rank_t give the ranks
// of the lane bits and matrix bits).  Items of a launch share k.
struct BatchTileItem {
     uint64_t state_off;
     uint64_t mat_off;
     uint64_t block_start;
     uint32_t hoff[32];
     uint8_t ins[5];  // the pass's positions, ascending
     uint8_t nins;
     uint8_t rank_n[6];
     uint8_t rank_t[5];
     uint8_t log_blocks;  // log2 of the item's blocks
     uint8_t pad[6];
};

// Waves per block of the batched tile kernel (LDS holds 64 * 2^k amplitudes
// per wave).
inline int batch_tile_wpb(int k, Prec prec) { return (64 << k) * static_cast<int>(amp_bytes(prec)) > 16384 ? 2 : 4; }
// Builds the tile item of a k-qubit pass at ascending positions pos over a
// state of 2^n amplitudes; false if the tile form does not apply.
bool make_tile_item(int n, int k, const uint8_t* pos, Prec prec, BatchTileItem& it);
hipError_t launch_batch_tile(void* arena, Prec prec, const void* mats, const BatchTileItem* items,
                             int nitems, int k, uint64_t total_blocks, hipStream_t s);

constexpr int kBatchLayerBlock = 256;
// A state of at most this many bytes stays in LDS for its whole circuit.
constexpr size_t kBatchLdsBytes = 128 * 1024;

// Persistent launch over the workgroup-owned circuits: workgroup w runs
// circuits wg_list[wg_off[w] .. wg_off[w+1]) from |0..0> to the end.
// kmax is the widest pass (<= 5); with the precision it picks the
// workgroup size (batch_wg_threads).
hipError_t launch_batch_wg(void* arena, Prec prec, const void* mats, const BatchPass* passes,
                           const BatchCirc* circs, const uint32_t* wg_off,
                           const uint32_t* wg_list, int n_wg, int kmax, hipStream_t s);
int batch_wg_threads(Prec prec, int kmax);

// One layered launch: every item's pass, total_blocks = sum of item blocks.
hipError_t launch_batch_layer(void* arena, Prec prec, const void* mats, const BatchItem* items,
                              int nitems, int k, uint64_t total_blocks, hipStream_t s);

// arena[offsets[c]] = 1 for c < ncirc.
hipError_t launch_set_ones(void* arena, Prec prec, const uint64_t* offsets, int ncirc,
                           hipStream_t s);

// ---- serving: launches over requests at different points of their circuits
// One pass each of many states, widths mixed (item.k <= kmax).
hipError_t launch_serve_layer(void* arena, Prec prec, const void* mats, const BatchItem* items,
                              int nitems, uint64_t total_blocks, int kmax, hipStream_t s);
// Amplitude blocks per item of launch_init_states.
constexpr int kInitAmpsPerBlock = 1024;
// |0..0> for each item's state: item {state_off, block_start, log_groups = n}.
hipError_t launch_init_states(void* arena, Prec prec, const BatchItem* items, int nitems,
                              uint64_t total_blocks, hipStream_t s);
// Probability of each block of 2^log_groups amplitudes: item {state_off,
// mat_off = first output slot, block_start, log_groups}; one workgroup per
// block; out[mat_off + j] = sum over block j of |psi|^2 (FP64).
hipError_t launch_block_sums_multi(const void* arena, Prec prec, const BatchItem* items,
                                   int nitems, uint64_t total_blocks, double* out, hipStream_t s);

// Sampling job: `shots` computational-basis outcomes of one state, shot s
// drawing u_s from Philox4x32-10 (counter s, key = seed).  sums[sums_off ..
// + 2^log_blocks) holds the block probabilities (launch_block_sums_multi);
// they are turned into a running sum in place.
struct SampleJob {
     uint64_t state_off;
     uint64_t sums_off;
     uint64_t out_off;
     uint64_t seed;
     uint32_t log_blocks;
     uint32_t log_block_amps;
     uint32_t shots;
     uint32_t pad;
};
// Probability blocks of a sample job over 2^n amplitudes: blocks of at
// least 256 amplitudes, at most 2^15 blocks.
inline uint32_t sample_log_blocks(int n)
{
     const int lb = n - 8;
     return static_cast<uint32_t>(lb < 0 ? 0 : (lb > 15 ? 15 : lb));
}
// One workgroup per job; out[out_off + s] = outcome of shot s.
hipError_t launch_sample(const void* arena, Prec prec, const SampleJob* jobs, int njobs,
                         double* sums, uint64_t* out, hipStream_t s);

// ---- tile programs: many fused ops in one HBM pass --------------------------
// A tile holds the 2^kTileBits amplitudes whose index bits outside
// V = {0..5} + hpos are fixed; tile rank r numbers the V bits in ascending
// order, so rank 0..5 are index bits 0..5 (64 contiguous amplitudes) and
// rank 6 + j is index bit hpos[j].  One workgroup loads its tile into LDS,
// applies every op in order (workgroup barrier between ops) and writes it
// back: one read and one write of the state for the whole op list.
constexpr int kTileBits = 12;
constexpr int kTileHigh = kTileBits - 6;
struct TileOp {
     uint64_t mat_off;  // first complex entry of the op's 2^k x 2^k matrix (tile_mat_slot layout)
     uint8_t k;
     uint8_t rpos[5];  // ascending tile ranks; matrix bit l <-> rpos[l]
     uint8_t pad[2];
};
// Where entry i = o * D + j (row o, column j, D = 2^k) of a tile op's matrix
// sits in the matrix array: row major for ops of 1-2 qubits; for 3-5 qubits
// in the order the matrix-core passes read their A operands, so that lane l
// of a wave reads entry s * 64 + l (k >= 4) at its s-th operand load, with
// no per-load address arithmetic and consecutive lanes on consecutive LDS
// banks:
//   k = 4, 5: (o, j) at ((o / 16) * D / 4 + j / 4) * 64 + (j % 4) * 16 + o % 16
//   k = 3:    (o, j) at (j / 4) * 32 + (j % 4) * 8 + o
// The tile kernels read every op matrix through the same map.
inline uint32_t tile_mat_slot(int k, uint32_t i)
{
     const uint32_t D = 1u << k, o = i >> k, j = i & (D - 1);
     if (k >= 4) return ((o >> 4) * (D >> 2) + (j >> 2)) * 64 + (j & 3) * 16 + (o & 15);
     if (k == 3) return (j >> 2) * 32 + (j & 3) * 8 + o;
     return i;
}
// Writes a tile op's row-major 2^k x 2^k matrix m into dst (2^(2k) entries
// of the state's precision) in the tile_mat_slot layout.
inline void put_tile_matrix(void* dst, const std::complex<double>* m, int k, Prec prec)
{
     const uint32_t n = 1u << (2 * k);
     for (uint32_t i = 0; i < n; ++i) {
          const uint32_t s = tile_mat_slot(k, i);
          if (prec == Prec::C128) {
               double* d = static_cast<double*>(dst) + 2 * s;
               d[0] = m[i].real();
               d[1] = m[i].imag();
          }
          else {
               float* d = static_cast<float*>(dst) + 2 * s;
               d[0] = static_cast<float>(m[i].real());
               d[1] = static_cast<float>(m[i].imag());
          }
     }
}
struct TileHigh {
     uint8_t hpos[kTileHigh];  // ascending index bits >= 6
};
// One tile pass of one circuit inside a layered launch of tile programs.
struct BatchTileProg {
     uint64_t state_off;
     uint64_t block_start;
     uint32_t op_first;  // first op in the launch's op array
     uint16_t nops;
     uint8_t log_blocks;  // log2 of the item's tiles (n - kTileBits)
     uint8_t hpos[kTileHigh];
     uint8_t pad[7];
};
hipError_t launch_batch_tprog(void* arena, Prec prec, const void* mats, const BatchTileProg* items,
                              int nitems, const TileOp* ops, int kmax, uint64_t total_blocks,
                              hipStream_t s);
// One tile pass over a state of 2^n amplitudes (n >= kTileBits): ops[0..nops)
// (device array) with matrices in `mats` (device, state precision).
hipError_t launch_tile_pass(void* psi, Prec prec, int n_qubits, const TileHigh& h, const TileOp* ops,
                            int nops, int kmax, const void* mats, hipStream_t s);

// ---- host side ------------------------------------------------------------
struct BatchGate {
     std::vector<std::complex<double>> m;  // 2^t x 2^t, row-major, bit l <-> targets[l]
     std::vector<int64_t> targets;
     std::vector<int64_t> ctrls;
};

// One fused pass of a planned circuit: 2^k x 2^k matrix (FP64, row-major)
// over ascending positions.
struct PlannedPass {
     int k = 0;
     uint8_t pos[5] = {0, 0, 0, 0, 0};
     std::vector<std::complex<double>> m;
};
// One tile pass of a planned circuit: its planned bits and its fused ops,
// whose positions are tile ranks (0..5 = index bits 0..5, 6 + j = hpos[j]).
struct PlannedTile {
     uint8_t hpos[kTileHigh];
     std::vector<PlannedPass> ops;
};
// Clusters of at most k qubits (lookahead packer), each fused into one dense
// matrix over its ascending qubits, controls folded in as diag(I, U).
std::vector<PlannedPass> plan_circuit(const std::vector<BatchGate>& gates, int k);
// Tile passes (plan_tile_passes) with ops of at most op_width qubits fused.
std::vector<PlannedTile> plan_circuit_tiles(const std::vector<BatchGate>& gates, int n, int op_width);
// One planned tile pass (gate indices into `gates`) with its ops fused.
PlannedTile fuse_tile(const std::vector<BatchGate>& gates, const TilePlan& tp);
// The gates' positions as the schedulers take them.
GateList gate_list(const std::vector<BatchGate>& gates);

struct BatchStats {
     uint64_t circuits = 0;
     uint64_t passes = 0;           // fused passes over all circuits
     uint64_t wg_circuits = 0;      // circuits owned by one workgroup
     uint64_t lds_circuits = 0;     // ... of which LDS-resident
     uint64_t layer_launches = 0;   // layered launches in the graph
     uint64_t arena_bytes = 0;
     uint64_t matrix_bytes = 0;
     uint64_t amp_passes = 0;       // sum over passes of 2^n (amplitudes touched)
     int wg_grid = 0;
     int kmax = 0;                  // widest fused pass
     uint64_t padded_launches = 0;  // layered launches whose 2-D grid has padding blocks
     uint64_t tile_passes = 0;      // HBM passes run as tile programs
     uint64_t tile_ops = 0;         // ops inside them
     uint64_t tile_mfma_flops = 0;  // 8 * 2^k * 2^n summed over their ops of k >= 3
};

class Batch {
public:
     Batch(Prec prec, int device);
     ~Batch();
     Batch(const Batch&) = delete;
     Batch& operator=(const Batch&) = delete;

     // Returns the circuit's index.  Qubits are 0..n-1; every gate id must be
     // below n and a gate's targets and controls distinct.
     int64_t add_circuit(int n, std::vector<BatchGate> gates);
     // Appends `count` copies of circuit c; returns the first copy's index.
     int64_t add_copies(int64_t c, int64_t count);
     // k_by_n[n] = fusion width for circuits of n qubits (0 = default 4);
     // circuits of more than wg_max_qubits qubits take the layered path, whose
     // passes may use the wave-tile form from tile_min_qubits qubits on;
     // from prog_min_qubits qubits on (at least 12) a layered circuit runs as
     // tile programs instead (ops of at most k_by_n[n] qubits, several per
     // HBM pass).
     void set_plan(const std::vector<int>& k_by_n, int wg_max_qubits, int tile_min_qubits,
                   int prog_min_qubits);
     // Plans and fuses every circuit and lays out the device descriptors on
     // the host, then checks every launch's blocks and addresses against the
     // arena, matrix and item bounds (throws on a violation).  Needs no device.
     void prepare();
     // prepare(), then uploads and captures the graph (replaces a previous build).
     void build();
     // Replays the graph and waits; returns the device time in ms.
     double run();
     void get_state(int64_t c, std::complex<double>* out) const;
     int num_qubits(int64_t c) const;
     int64_t num_circuits() const { return static_cast<int64_t>(circ_.size()); }
     const BatchStats& stats() const { return stats_; }
     int fusion_width(int n) const;

     // Persistent workgroups used when no device is queried (prepare()).
     static constexpr int kDefaultWorkgroups = 256;

private:
     struct Circ {
          int n;
          std::vector<BatchGate> gates;
          uint64_t state_off = 0;
     };
     enum LaunchKind : int { kDirect = 0, kTileForm = 1, kTileProgram = 2 };
     struct LayerLaunch {
          int k;  // fused width, or the widest op of a tile-program launch
          int kind;
          size_t first, count;
          uint64_t blocks;
     };
     // Host image of a build: everything the device needs.
     struct HostPlan {
          std::vector<char> mats;
          uint64_t mat_entries = 0;
          uint64_t arena_amps = 0, layer_amps = 0;
          std::vector<BatchPass> pass;
          std::vector<BatchCirc> circ;
          std::vector<uint32_t> wg_off, wg_list;
          int n_wg = 0, kmax_wg = 1;
          std::vector<BatchItem> items;
          std::vector<BatchTileItem> titems;
          std::vector<BatchTileProg> progs;  // tile-program items
          std::vector<TileOp> tops;          // their ops
          std::vector<int> item_n, titem_n, prog_n;  // qubits of each item's circuit (checks)
          std::vector<LayerLaunch> launches;
          std::vector<uint64_t> ones;
     };
     void ensure_device();
     void release_device();
     HostPlan plan_host(int workgroups);
     void check_plan(const HostPlan& hp) const;

     Prec prec_;
     int device_;
     hipStream_t s1_ = nullptr, s2_ = nullptr;
     hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr, ev_t0_ = nullptr, ev_t1_ = nullptr;
     hipGraph_t graph_ = nullptr;
     hipGraphExec_t exec_ = nullptr;
     std::vector<Circ> circ_;
     std::vector<int> k_by_n_;
     int wg_max_qubits_ = 16;
     int tile_min_qubits_ = 0;
     int prog_min_qubits_ = 31;
     bool built_ = false;
     bool device_ready_ = false;
     void* d_arena_ = nullptr;
     void* d_mats_ = nullptr;
     void* d_desc_ = nullptr;  // passes, circs, wg lists, items, one-offsets
     BatchStats stats_;
};

}  // namespace hq
