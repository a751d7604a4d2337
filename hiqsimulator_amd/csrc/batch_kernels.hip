// gfx950 kernels of the batched many-circuit path (batch.hpp).
//
// Each fused pass is the dense k-qubit update of the single-state engine
// (kernels/intrin/kernel{1..5}.hpp of the reference: for every base I with
// the target bits 0, v_j = psi[I + off_j], psi[I + off_i] = sum_j m[i][j] v_j),
// applied to many states:
//  * k_batch_wg: persistent; each workgroup owns whole circuits.  A state of
//    at most 128 KiB is initialised to |0..0> in LDS, every pass of the
//    circuit runs on LDS with a workgroup barrier between passes, and the
//    final state is written to HBM once.  A larger state is streamed through
//    L2/HBM by its workgroup in the same way.  No inter-workgroup
//    synchronisation exists, so every workgroup ends after its own list.
//  * k_batch_layer: one pass each of many large circuits; blocks map to
//    items by binary search over the items' first blocks.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "batch.hpp"
#include "device_common.hpp"

namespace hq {
namespace {

template <typename V>
__device__ __forceinline__ V basis_amp(bool one)
{
     V a;
     a.x = one ? 1 : 0;
     a.y = 0;
     return a;
}

// Item lookup of a block-mapped launch: the last item whose first block is
// at or below b.
template <typename Item>
__device__ __forceinline__ int find_item(const Item* __restrict__ items, int nitems, uint64_t b)
{
     int lo = 0, hi = nitems - 1;
     while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (items[mid].block_start <= b) lo = mid;
          else hi = mid - 1;
     }
     return lo;
}

// One fused pass of one workgroup over a state of 2^n amplitudes at psi
// (LDS or global memory); mat is the pass's matrix in LDS, pos ascending.
// The rows are not unrolled: a row is read (uniform address, broadcast),
// reduced against the group held in registers and stored before the next.
template <int K, typename R, int NT>
__device__ __forceinline__ void wg_pass(VT<R>* psi, uint32_t n, const VT<R>* mat,
                                        const uint8_t (&pos)[5])
{
     using V = VT<R>;
     constexpr int D = 1 << K;
     const uint32_t ng = 1u << (n - K);
     uint32_t off[D];
#pragma unroll
     for (int j = 0; j < D; ++j) {
          uint32_t o = 0;
#pragma unroll
          for (int l = 0; l < K; ++l)
               if ((j >> l) & 1) o |= 1u << pos[l];
          off[j] = o;
     }
     for (uint32_t t = threadIdx.x; t < ng; t += NT) {
          uint32_t b = t;
#pragma unroll
          for (int l = 0; l < K; ++l) {
               const uint32_t lo = b & ((1u << pos[l]) - 1u);
               b = ((b ^ lo) << 1) | lo;
          }
          V v[D];
#pragma unroll
          for (int j = 0; j < D; ++j) v[j] = psi[b + off[j]];
#pragma unroll 1
          for (int i = 0; i < D; ++i) {
               const V* const row = mat + i * D;
               R re = 0, im = 0;
#pragma unroll
               for (int j = 0; j < D; ++j) {
                    const V m = row[j];
                    re = fma(m.x, v[j].x, re);
                    re = fma(-m.y, v[j].y, re);
                    im = fma(m.x, v[j].y, im);
                    im = fma(m.y, v[j].x, im);
               }
               uint32_t oi = 0;
#pragma unroll
               for (int l = 0; l < K; ++l)
                    if ((i >> l) & 1) oi |= 1u << pos[l];
               V o;
               o.x = re;
               o.y = im;
               psi[b + oi] = o;
          }
     }
}

// Offset of amplitude j of a group: bit l of j deposited at pos[l].
template <int K>
__device__ __forceinline__ uint32_t amp_offset(int j, const uint8_t (&pos)[5])
{
     uint32_t o = 0;
#pragma unroll
     for (int l = 0; l < K; ++l)
          if ((j >> l) & 1) o |= 1u << pos[l];
     return o;
}

// LDS slot of element r of a tile program's tile: the low four index bits
// XOR-ed with bits 4..7 (a bijection of every aligned 256-element block).
// LDS has 64 banks of 4 bytes, serving a 64-lane 8-byte access in two
// 32-lane halves and a 16-byte access in four 16-lane quarters.  In the
// identity layout the matrix-core operand loads put lanes 2^rank apart on
// one bank (up to 16 accesses per bank in complex128); with this layout,
// modelled over the ops of the 4096-circuit batch, the tile accesses need
// 0.79 (complex64) / 0.54 (complex128) of the bank cycles.  (The staged op
// matrices need no swizzle: their layout, mat_slot, puts consecutive lanes
// on consecutive entries.)  It is linear over XOR, so for a group base b and
// an amplitude offset o with disjoint bits, tslot(b + o) = tslot(b) ^ tslot(o).
__device__ __forceinline__ uint32_t tslot(uint32_t r)
{
     return r ^ ((r >> 4) & 15u);
}

// Element of an LDS array at a byte offset (offsets are combined with XOR in
// bytes, so no per-access shift of an element index is needed).
template <typename V>
__device__ __forceinline__ V& lds_at(V* base, uint32_t bytes)
{
     return *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + bytes);
}

// The compiler folds (a << s) ^ (b << s) back into (a ^ b) << s, which puts
// a shift in front of every LDS access; an empty asm statement makes a
// byte offset opaque to it (no instruction is emitted).
__device__ __forceinline__ uint32_t opaque_v(uint32_t x)
{
     asm("" : "+v"(x));
     return x;
}
__device__ __forceinline__ uint32_t opaque_s(uint32_t x)
{
     asm("" : "+s"(x));
     return x;
}

// Entry (row o, column j) of a tile op's 2^K x 2^K matrix, in global memory
// and in its LDS copy alike: row major for K <= 2; for K >= 3 the order in
// which the matrix-core passes read their A operands (the host writes it so,
// tile_mat_slot), where lane (q, c) = (l / 16, l % 16) reads M[16 rb + c][4 t + q]
// at entry (rb D/4 + t) 64 + l (K >= 4) and M[o][4 t + q] at t 32 + q 8 + o
// (K = 3).
template <int K>
__device__ __forceinline__ uint32_t mat_slot(uint32_t o, uint32_t j)
{
     constexpr uint32_t D = 1u << K;
     if constexpr (K >= 4) return ((o >> 4) * (D >> 2) + (j >> 2)) * 64 + (j & 3) * 16 + (o & 15);
     else if constexpr (K == 3) return (j >> 2) * 32 + (j & 3) * 8 + o;
     else return o * D + j;
}

// Group index v spread over the non-target positions: a zero bit inserted
// at each of pos[0] < ... < pos[K-1].  Bitwise, so index parts with disjoint
// bits spread to disjoint bits: the addresses below split into a per-lane
// part, computed once per op, and a part uniform across the wave, which
// the compiler keeps in scalar registers (the scalar unit computes it, the
// vector unit only XORs it in).
template <int K>
__device__ __forceinline__ uint32_t insert_targets(uint32_t v, const uint8_t (&pos)[5])
{
#pragma unroll
     for (int l = 0; l < K; ++l) {
          const uint32_t lo = v & ((1u << pos[l]) - 1u);
          v = ((v ^ lo) << 1) | lo;
     }
     return v;
}

// The same pass on a 2^N-amplitude LDS tile with the matrix read from
// global memory at a uniform address (scalar loads into SGPRs, shared by the
// whole wave; entries at mat_slot), two rows at a time: no LDS traffic for
// the matrix, and no SGPR spills (a fully unrolled row loop would hold the
// whole matrix).  When
// the groups outnumber the threads, each thread reads all of its groups
// before its first write, so their LDS latencies overlap.
template <int K, typename R, int NTH, int GN, int N = kTileBits>
__device__ __forceinline__ void gm_groups(VT<R>* psi, const VT<R>* __restrict__ mat, const uint8_t (&pos)[5],
                                          const uint32_t g0, const uint32_t t0)
{
     // thread t0 of NTH takes groups g0 + t0 + NTH * g of the GN from g0
     using V = VT<R>;
     constexpr int D = 1 << K;
     constexpr int G = GN >= NTH ? GN / NTH : 1;  // groups per thread
     static_assert(GN < NTH || GN % NTH == 0, "groups must divide over the threads");
     static_assert((NTH & (NTH - 1)) == 0, "NTH must be a power of two");
     constexpr uint32_t SH = sizeof(V) == 16 ? 4 : 3;  // byte offsets below
     uint32_t off[D];  // LDS slots of the amplitude offsets (tslot; uniform)
#pragma unroll
     for (int j = 0; j < D; ++j) off[j] = opaque_s(tslot(amp_offset<K>(j, pos)) << SH);
     if (GN < NTH && t0 >= static_cast<uint32_t>(GN)) return;
     // group g0 + NTH * g + t0 with g0 a multiple of NTH and t0 < NTH: the
     // lane part t0 and the uniform part have disjoint bits
     const uint32_t vt = opaque_v(tslot(insert_targets<K>(t0, pos)) << SH);
     uint32_t base[G];
     V v[G][D];
#pragma unroll
     for (int g = 0; g < G; ++g) {
          base[g] = vt ^ opaque_s(tslot(insert_targets<K>(g0 + static_cast<uint32_t>(g) * NTH, pos)) << SH);
#pragma unroll
          for (int j = 0; j < D; ++j) v[g][j] = lds_at(psi, base[g] ^ off[j]);
     }
#pragma unroll 2
     for (int i = 0; i < D; ++i) {
#pragma unroll
          for (int g = 0; g < G; ++g) {
               R re = 0, im = 0;
#pragma unroll
               for (int j = 0; j < D; ++j) {
                    const V m = mat[mat_slot<K>(static_cast<uint32_t>(i), static_cast<uint32_t>(j))];
                    re = fma(m.x, v[g][j].x, re);
                    re = fma(-m.y, v[g][j].y, re);
                    im = fma(m.x, v[g][j].y, im);
                    im = fma(m.y, v[g][j].x, im);
               }
               V o;
               o.x = re;
               o.y = im;
               lds_at(psi, base[g] ^ off[i]) = o;
          }
     }
}

template <int K, typename R, int NT, int N = kTileBits>
__device__ __forceinline__ void wg_pass_gm(VT<R>* psi, const VT<R>* __restrict__ mat, const uint8_t (&pos)[5])
{
     gm_groups<K, R, NT, (1 << (N - K)), N>(psi, mat, pos, 0, threadIdx.x);
}

// The same pass on the matrix cores, for 3 <= K <= 5.  The op is a real
// product Out = A * In with In the 2D x (ng) matrix of the groups'
// amplitudes (real and imaginary parts as separate rows) and A the 2D x 2D
// real form of the complex matrix M:
//   out.re = sum_j  Re M_oj x_j.re - Im M_oj x_j.im
//   out.im = sum_j  Im M_oj x_j.re + Re M_oj x_j.im.
// It runs as 16x16x4 MFMA tiles: 16 real output rows x 16 groups, 4 real
// input rows per instruction.  The order of the real rows inside a tile is
// chosen so that every lane moves whole complex amplitudes:
//   * input: instruction s, lane (q = lane/16, c = lane%16) supplies part
//     s%2 of amplitude 4*(s/2) + q of group c, so one LDS load of a complex
//     amplitude feeds two instructions;
//   * output: lane (q, c) receives four rows of its column (rows 4q..4q+3
//     for the f32 instruction, rows q, q+4, q+8, q+12 for the f64 one, as
//     measured on gfx950); the real rows are numbered so that these four are
//     the real and imaginary parts of amplitudes 8*rb + 2q and 8*rb + 2q + 1
//     of group c (rb = the 16-row block).
// Every wave owns whole 16-group column blocks and reads all of a block's
// amplitudes before its first write, so the pass stays in place without a
// barrier of its own.
template <typename R>
struct Mfma16;
template <>
struct Mfma16<double> {
     typedef double acc_t __attribute__((ext_vector_type(4)));
     static __device__ __forceinline__ acc_t step(double a, double b, acc_t c)
     {
          return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
     }
     // real row r (0..15) of a block: amplitude offset in the block, imaginary part?
     static __device__ __forceinline__ int row_amp(int r) { return 2 * (r & 3) + (r >> 3); }
     static __device__ __forceinline__ bool row_im(int r) { return (r >> 2) & 1; }
     // output row of accumulator element j on lanes 16q..16q+15
     static __device__ __forceinline__ int out_row(int q, int j) { return q + 4 * j; }
};
template <>
struct Mfma16<float> {
     typedef float acc_t __attribute__((ext_vector_type(4)));
     static __device__ __forceinline__ acc_t step(float a, float b, acc_t c)
     {
          return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
     }
     static __device__ __forceinline__ int row_amp(int r) { return r >> 1; }
     static __device__ __forceinline__ bool row_im(int r) { return r & 1; }
     static __device__ __forceinline__ int out_row(int q, int j) { return 4 * q + j; }
};


// mat[moff ..] is the op's matrix (mat_slot layout): in LDS when SW (staged),
// else in global memory.
template <int K, typename R, int NWV, int NCB, bool SW, int N = kTileBits>
__device__ __forceinline__ void mfma_blocks(VT<R>* psi, const VT<R>* __restrict__ mat, const uint32_t moff,
                                            const uint8_t (&pos)[5], const int w, const uint32_t cb0)
{
     // wave w of NWV takes column blocks cb0 + w + NWV * i of the NCB from cb0
     static_assert(K >= 3 && K <= 5, "MFMA pass needs 3 <= K <= 5");
     using V = VT<R>;
     using M = Mfma16<R>;
     using acc_t = typename M::acc_t;
     constexpr int D = 1 << K;
     constexpr int RB = D / 8;  // 16-row blocks of the real output
     constexpr int T = D / 4;   // complex input amplitudes per lane
     static_assert(NCB % NWV == 0, "column blocks must divide over the waves");
     constexpr int CB = NCB / NWV;  // column blocks per wave
     constexpr bool kHoldA = RB * T <= 8;
     const int lane = threadIdx.x & 63;
     const int q = lane >> 4, c = lane & 15;
     // A fragments: lane (q, c) holds row c of block rb (output amplitude
     // 8*rb + row_amp(c), part row_im(c)) at input amplitudes 4t + q.
     auto load_a = [&](int rb, R (&a)[2 * T]) {
          const int o = rb * 8 + M::row_amp(c);
          const bool im_row = M::row_im(c);
#pragma unroll
          for (int t = 0; t < T; ++t) {
               const V m = mat[moff + mat_slot<K>(static_cast<uint32_t>(o), static_cast<uint32_t>(4 * t + q))];
               a[2 * t] = im_row ? m.y : m.x;       // times x.re
               a[2 * t + 1] = im_row ? m.x : -m.y;  // times x.im
          }
     };
     R ahold[kHoldA ? RB : 1][2 * T];
     if constexpr (kHoldA) {
#pragma unroll
          for (int rb = 0; rb < RB; ++rb) load_a(rb, ahold[rb]);
     }
     // LDS slots split into lane and uniform parts (disjoint index bits):
     // group c of a column block, input amplitude 4t + q, output amplitudes
     // 8 rb + 2q + h
     constexpr uint32_t SH = sizeof(V) == 16 ? 4 : 3;  // byte offsets below
     const uint32_t vc = tslot(insert_targets<K>(static_cast<uint32_t>(c), pos));
     const uint32_t vin = opaque_v((vc ^ tslot(amp_offset<K>(q, pos))) << SH);
     const uint32_t vout = opaque_v((vc ^ tslot(amp_offset<K>(2 * q, pos))) << SH);
     uint32_t in_off[T], out_off[RB][2];  // uniform
#pragma unroll
     for (int t = 0; t < T; ++t) in_off[t] = opaque_s(tslot(amp_offset<K>(4 * t, pos)) << SH);
#pragma unroll
     for (int rb = 0; rb < RB; ++rb) {
          out_off[rb][0] = opaque_s(tslot(amp_offset<K>(rb * 8, pos)) << SH);
          out_off[rb][1] = opaque_s(tslot(amp_offset<K>(rb * 8 + 1, pos)) << SH);
     }
     // The wave's column blocks in groups of HB: every amplitude of a group
     // is read before the group's first write.
     constexpr int HB = CB < 2 ? CB : 2;
#pragma unroll
     for (int i0 = 0; i0 < CB; i0 += HB) {
          uint32_t base[HB];  // uniform part of the column block's groups
          V x[HB][T];
#pragma unroll
          for (int i = 0; i < HB; ++i) {
               base[i] = opaque_s(tslot(insert_targets<K>((cb0 + static_cast<uint32_t>(w + (i0 + i) * NWV)) * 16, pos)) << SH);
               const uint32_t bi = vin ^ base[i];
#pragma unroll
               for (int t = 0; t < T; ++t) x[i][t] = lds_at(psi, bi ^ in_off[t]);
          }
#pragma unroll
          for (int i = 0; i < HB; ++i) {
               const uint32_t bo = vout ^ base[i];
               auto block = [&](const R (&a)[2 * T], int rb) {
                    acc_t acc = {0, 0, 0, 0};
#pragma unroll
                    for (int t = 0; t < T; ++t) {
                         acc = M::step(a[2 * t], x[i][t].x, acc);
                         acc = M::step(a[2 * t + 1], x[i][t].y, acc);
                    }
                    V y0, y1;
                    y0.x = acc[0];
                    y0.y = acc[1];
                    y1.x = acc[2];
                    y1.y = acc[3];
                    lds_at(psi, bo ^ out_off[rb][0]) = y0;
                    lds_at(psi, bo ^ out_off[rb][1]) = y1;
               };
#pragma unroll
               for (int rb = 0; rb < RB; ++rb) {
                    if constexpr (kHoldA) {
                         block(ahold[rb], rb);
                    }
                    else {
                         R a[2 * T];
                         load_a(rb, a);
                         block(a, rb);
                    }
               }
          }
     }
}

// The same pass for 4 <= K <= 5 with three real D x D products instead of
// the 2D x 2D real form (a three-multiplication complex product in the
// manner of Gauss's; with M = A + iB and x = u + iv):
//   Q1 = A . (u + v),  Q2 = (A + B) . v,  Q3 = (B - A) . u,
//   out.re = Q1 - Q2 = A u - B v,  out.im = Q1 + Q3 = A v + B u,
// 3 D^2 / 64 instructions per 16 groups instead of 4 D^2 / 64, and two
// vector instructions per output amplitude (the form P1 = A u, P2 = B v,
// P3 = (A + B)(u + v), out.im = P3 - P1 - P2 takes three).  Here the
// real rows of a 16-row block are the amplitudes 16 * rb + r in natural
// order; instruction s takes input amplitude 4 * s + q from lane (q, c), so
// one LDS load of a complex amplitude feeds three instructions, and lane
// (q, c)'s accumulator element j is output amplitude 16 * rb + out_row(q, j)
// of group c.  (For K = 3, D = 8 rows would fill half of each 16-row
// instruction, so those ops keep the 2D x 2D form.)  The rounding error
// grows from ~eps |M| |x| to ~2 eps (|A| + |B|)(|u| + |v|).
template <int K, typename R, int NWV, int NCB, bool SW, int N = kTileBits>
__device__ __forceinline__ void mfma_blocks_3m(VT<R>* psi, const VT<R>* __restrict__ mat, const uint32_t moff,
                                               const uint8_t (&pos)[5], const int w, const uint32_t cb0)
{
     static_assert(K >= 4 && K <= 5, "three-product MFMA pass needs 4 <= K <= 5");
     using V = VT<R>;
     using M = Mfma16<R>;
     using acc_t = typename M::acc_t;
     constexpr int D = 1 << K;
     constexpr int RB = D / 16;  // 16-row blocks of each product
     constexpr int T = D / 4;    // instructions per product and input amplitudes per lane
     static_assert(NCB % NWV == 0, "column blocks must divide over the waves");
     constexpr int CB = NCB / NWV;
     constexpr bool kHoldA = RB == 1;  // K = 4: 3 T values per lane
     const int lane = threadIdx.x & 63;
     const int q = lane >> 4, c = lane & 15;
     // A fragments of row block rb: lane (q, c) holds Re, Im and Re + Im of
     // M[16 rb + c][4 t + q], at entry (rb T + t) 64 + lane (mat_slot), so
     // the loads differ only by constant offsets.
     const V* const alane = mat + moff + lane;
     auto load_a = [&](int rb, R (&a)[3 * T]) {
#pragma unroll
          for (int t = 0; t < T; ++t) {
               const V m = alane[(rb * T + t) * 64];
               a[3 * t] = m.x;               // A, times u + v
               a[3 * t + 1] = m.x + m.y;     // A + B, times v
               a[3 * t + 2] = m.y - m.x;     // B - A, times u
          }
     };
     R ahold[3 * T];
     if constexpr (kHoldA) load_a(0, ahold);
     // LDS slots split into lane and uniform parts (disjoint index bits):
     // group c of a column block, input amplitude 4t + q, output amplitude
     // 16 rb + out_row(q, j) = 16 rb + out_row(q, 0) ^ out_row(0, j)
     constexpr uint32_t SH = sizeof(V) == 16 ? 4 : 3;  // byte offsets below
     const uint32_t vc = tslot(insert_targets<K>(static_cast<uint32_t>(c), pos));
     const uint32_t vin = opaque_v((vc ^ tslot(amp_offset<K>(q, pos))) << SH);
     const uint32_t vout = opaque_v((vc ^ tslot(amp_offset<K>(M::out_row(q, 0), pos))) << SH);
     uint32_t in_off[T], out_off[RB][4];  // uniform
#pragma unroll
     for (int t = 0; t < T; ++t) in_off[t] = opaque_s(tslot(amp_offset<K>(4 * t, pos)) << SH);
#pragma unroll
     for (int rb = 0; rb < RB; ++rb)
#pragma unroll
          for (int j = 0; j < 4; ++j)
               out_off[rb][j] = opaque_s(tslot(amp_offset<K>(rb * 16 + M::out_row(0, j), pos)) << SH);
     constexpr int HB = CB < 2 ? CB : 2;
#pragma unroll
     for (int i0 = 0; i0 < CB; i0 += HB) {
          uint32_t base[HB];  // uniform part of the column block's groups
          V x[HB][T];
#pragma unroll
          for (int i = 0; i < HB; ++i) {
               base[i] = opaque_s(tslot(insert_targets<K>((cb0 + static_cast<uint32_t>(w + (i0 + i) * NWV)) * 16, pos)) << SH);
               const uint32_t bi = vin ^ base[i];
#pragma unroll
               for (int t = 0; t < T; ++t) x[i][t] = lds_at(psi, bi ^ in_off[t]);
          }
#pragma unroll
          for (int i = 0; i < HB; ++i) {
               const uint32_t bo = vout ^ base[i];
               auto block = [&](const R (&a)[3 * T], int rb) {
                    acc_t p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
#pragma unroll
                    for (int t = 0; t < T; ++t) {
                         p1 = M::step(a[3 * t], x[i][t].x + x[i][t].y, p1);
                         p2 = M::step(a[3 * t + 1], x[i][t].y, p2);
                         p3 = M::step(a[3 * t + 2], x[i][t].x, p3);
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                         V y;
                         y.x = p1[j] - p2[j];
                         y.y = p1[j] + p3[j];
                         lds_at(psi, bo ^ out_off[rb][j]) = y;
                    }
               };
#pragma unroll
               for (int rb = 0; rb < RB; ++rb) {
                    if constexpr (kHoldA) {
                         block(ahold, rb);
                    }
                    else {
                         // one row block at a time (K = 5): interleaving the
                         // two measured 2.5% slower on the 5-qubit one-op
                         // passes (the operand loads were then issued just
                         // before their use)
                         R a[3 * T];
                         load_a(rb, a);
                         block(a, rb);
                         __builtin_amdgcn_sched_barrier(0);
                    }
               }
          }
     }
}

template <int K, typename R, int NT, bool SW, int N = kTileBits>
__device__ __forceinline__ void wg_pass_mfma(VT<R>* psi, const VT<R>* __restrict__ mat, const uint32_t moff,
                                             const uint8_t (&pos)[5])
{
     const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
     if constexpr (K >= 4) mfma_blocks_3m<K, R, NT / 64, (1 << (N - K)) / 16, SW, N>(psi, mat, moff, pos, w, 0);
     else mfma_blocks<K, R, NT / 64, (1 << (N - K)) / 16, SW, N>(psi, mat, moff, pos, w, 0);
}

// |0..0> followed by every pass of circuit c, on psi.  Each pass's matrix is
// staged in LDS (smat) and read from there with uniform-address loads, which
// keeps the scalar registers free for the group offsets.
template <typename R, int KMAX, int NT>
__device__ __forceinline__ void wg_circuit(VT<R>* psi, VT<R>* smat, const BatchCirc& c,
                                           const BatchPass* __restrict__ passes,
                                           const VT<R>* __restrict__ mats)
{
     using V = VT<R>;
     const uint32_t N = 1u << c.n;
     for (uint32_t i = threadIdx.x; i < N; i += NT) psi[i] = basis_amp<V>(i == 0);
     for (uint32_t p = c.pass_begin; p < c.pass_end; ++p) {
          const BatchPass ps = passes[p];
          const uint32_t dd = 1u << (2 * ps.k);
          for (uint32_t i = threadIdx.x; i < dd; i += NT) smat[i] = mats[ps.mat_off + i];
          __syncthreads();
          switch (ps.k) {
               case 1: wg_pass<1, R, NT>(psi, c.n, smat, ps.pos); break;
               case 2: wg_pass<2, R, NT>(psi, c.n, smat, ps.pos); break;
               case 3: wg_pass<3, R, NT>(psi, c.n, smat, ps.pos); break;
               case 4: wg_pass<4, R, NT>(psi, c.n, smat, ps.pos); break;
               default:
                    if constexpr (KMAX >= 5) wg_pass<5, R, NT>(psi, c.n, smat, ps.pos);
                    break;
          }
          __syncthreads();
     }
}

template <typename R, int KMAX, int NT>
__global__ void __launch_bounds__(NT)
    k_batch_wg(VT<R>* __restrict__ arena, const VT<R>* __restrict__ mats,
               const BatchPass* __restrict__ passes, const BatchCirc* __restrict__ circs,
               const uint32_t* __restrict__ wg_off, const uint32_t* __restrict__ wg_list)
{
     using V = VT<R>;
     constexpr uint32_t kLds = kBatchLdsBytes / sizeof(V);
     __shared__ V lds[kLds];
     __shared__ V smat[1 << (2 * KMAX)];
     const uint32_t c0 = wg_off[blockIdx.x], c1 = wg_off[blockIdx.x + 1];
     for (uint32_t ci = c0; ci < c1; ++ci) {
          const BatchCirc c = circs[wg_list[ci]];
          const uint32_t N = 1u << c.n;
          V* const g = arena + c.state_off;
          if (N <= kLds) {
               wg_circuit<R, KMAX, NT>(lds, smat, c, passes, mats);
               for (uint32_t i = threadIdx.x; i < N; i += NT) st_nt(g + i, lds[i]);
               __syncthreads();
          }
          else {
               wg_circuit<R, KMAX, NT>(g, smat, c, passes, mats);
          }
     }
}

// Tile programs.  A tile is the 2^kTileBits amplitudes whose index bits
// outside {0..5} u hpos are fixed: 64 chunks of 64 contiguous amplitudes,
// chunk c at base + pdep(c, hpos).  One workgroup loads its tile into LDS,
// runs the program's ops there (a workgroup barrier before each) and writes
// the tile back; with two (complex128) or five (complex64) workgroups per
// CU, one workgroup's ops overlap the others' HBM traffic (measured on a
// 30-qubit state: up to two 4-qubit or four 3-qubit ops per tile cost no
// time over the bare copy).
//
// Ops of at least MK qubits run on the matrix cores (wg_pass_mfma), the
// others on the vector units with the matrix read by uniform scalar loads
// (MK = 6: all on the vector units).  In complex128 the matrix-core ops read
// their matrices from an LDS copy (smat) loaded together with the tile, so
// the op phase issues no global loads; complex64 programs of at most
// 4-qubit ops read them from global memory and keep the LDS for a fifth
// workgroup.
// A descriptor read as 32-bit words at a wave-uniform address: the words
// come through scalar loads (a byte field read on its own would be a vector
// load, whose wait would also wait for the tile loads in flight).
template <typename T>
__device__ __forceinline__ T load_uniform(const T* __restrict__ p)
{
     static_assert(sizeof(T) % 4 == 0, "descriptor size must be a multiple of 4");
     constexpr int N = sizeof(T) / 4;
     const uint32_t* const q = reinterpret_cast<const uint32_t*>(p);
     uint32_t w[N];
#pragma unroll
     for (int i = 0; i < N; ++i) w[i] = q[i];
     T r;
     __builtin_memcpy(&r, w, sizeof(T));
     return r;
}

template <typename R, int NT>
struct TileIo {
     using V = VT<R>;
     static constexpr int NW = NT / 64;
     static constexpr int PER = (64 / NW);  // chunks per wave

     static __device__ __forceinline__ int wave() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
     static __device__ __forceinline__ uint64_t base_of(uint64_t tix, const uint8_t (&hpos)[kTileHigh])
     {
          uint64_t base = tix << 6;
#pragma unroll
          for (int j = 0; j < kTileHigh; ++j) {
               const uint64_t lo = base & ((1ull << hpos[j]) - 1ull);
               base = ((base ^ lo) << 1) | lo;
          }
          return base;
     }
     static __device__ __forceinline__ uint64_t chunk_off(int c, const uint8_t (&hpos)[kTileHigh])
     {
          uint64_t off = 0;
#pragma unroll
          for (int j = 0; j < kTileHigh; ++j) off |= static_cast<uint64_t>((c >> j) & 1) << hpos[j];
          return off;
     }
     static __device__ __forceinline__ void load(const V* __restrict__ psi, uint64_t tix,
                                                 const uint8_t (&hpos)[kTileHigh], V (&r)[PER])
     {
          const int lane = threadIdx.x & 63, w = wave();
          const V* const p = psi + base_of(tix, hpos) + lane;
#pragma unroll
          for (int i = 0; i < PER; ++i) r[i] = ld_nt(p + chunk_off(w + i * NW, hpos));
     }
     static __device__ __forceinline__ void store(V* __restrict__ psi, uint64_t tix,
                                                  const uint8_t (&hpos)[kTileHigh], const V (&r)[PER])
     {
          const int lane = threadIdx.x & 63, w = wave();
          V* const p = psi + base_of(tix, hpos) + lane;
#pragma unroll
          for (int i = 0; i < PER; ++i) st_nt(p + chunk_off(w + i * NW, hpos), r[i]);
     }
     static __device__ __forceinline__ void to_lds(V* tile, const V (&r)[PER])
     {
          const int lane = threadIdx.x & 63, w = wave();
#pragma unroll
          for (int i = 0; i < PER; ++i) tile[tslot(static_cast<uint32_t>((w + i * NW) * 64 + lane))] = r[i];
     }
     static __device__ __forceinline__ void from_lds(const V* tile, V (&r)[PER])
     {
          const int lane = threadIdx.x & 63, w = wave();
#pragma unroll
          for (int i = 0; i < PER; ++i) r[i] = tile[tslot(static_cast<uint32_t>((w + i * NW) * 64 + lane))];
     }
};

// LDS copy of the matrices of one tile program (matrix-core variants): the
// ops' matrices back to back in op order, as far as kStage elements go, in
// their global layout (mat_slot: the operand loads of consecutive lanes hit
// consecutive banks); ops beyond that read global memory.
template <typename R, int KMAX>
constexpr int tile_stage_elems()
{
     // 16 KiB complex128.  Complex64 programs of at most 4-qubit ops stage
     // nothing: without the 8 KiB their 32 KiB tile lets five workgroups
     // share a CU instead of four.  With 5-qubit ops the registers allow
     // four at most, and the 8 KiB copy is kept (one-op 5-qubit passes on
     // 30 qubits: 3.12-3.21 ms staged, 3.44-3.46 ms not).
     return sizeof(R) == 8 || KMAX >= 5 ? 1024 : 0;
}

// The LDS array for the staged matrices (none when N is 0).
template <typename V, int N>
__device__ __forceinline__ V* smat_buf()
{
     if constexpr (N > 0) {
          __shared__ V buf[N];
          return buf;
     }
     else {
          return nullptr;
     }
}

template <typename R, int NT, int KMAX>
struct MatStage {
     using V = VT<R>;
     static constexpr int ME = tile_stage_elems<R, KMAX>();
     static constexpr int PER = ME / NT > 0 ? ME / NT : 1;

     // Loads the matrices of ops[0..nops) that fit into r; returns the
     // number of staged elements.  Addresses first (scalar walk over the
     // ops, per-lane selects), then one unconditional load per register, so
     // no load waits for another.
     static __device__ __forceinline__ int load(V (&r)[PER], const V* __restrict__ mats,
                                                const TileOp* __restrict__ ops, int nops)
     {
          uint64_t addr[PER];
#pragma unroll
          for (int i = 0; i < PER; ++i) addr[i] = 0;
          int start = 0;
          for (int p = 0; p < nops; ++p) {
               const TileOp o = load_uniform(ops + p);
               const int sz = 1 << (2 * o.k);
               if (start + sz > ME) break;
#pragma unroll
               for (int i = 0; i < PER; ++i) {
                    const int e = static_cast<int>(threadIdx.x) + i * NT;
                    if (e >= start && e < start + sz) addr[i] = o.mat_off + static_cast<uint64_t>(e - start);
               }
               start += sz;
          }
          // nothing staged (no ops): mats may be empty, read its first entry never
          const V* const src = start > 0 ? mats : nullptr;
#pragma unroll
          for (int i = 0; i < PER; ++i) {
               V x{};
               if (src) x = src[addr[i]];
               r[i] = x;
          }
          return start;
     }
     static __device__ __forceinline__ void to_lds(V* smat, const V (&r)[PER], int n)
     {
#pragma unroll
          for (int i = 0; i < PER; ++i) {
               const int e = static_cast<int>(threadIdx.x) + i * NT;
               if (e < n) smat[e] = r[i];
          }
     }
};

// Op engines of the tile-program kernels: every op on the vector units, or
// ops of 3+ qubits on the matrix cores.  (Splitting the 3-qubit ops'
// groups between matrix-core waves and vector waves, which have the same
// FP64 peak, measured 5-25% slower than the matrix cores alone.)
constexpr int kTileVector = 0, kTileMatrix = 1;
constexpr int tile_mfma_min_k(int mode) { return mode == kTileVector ? 6 : 3; }

// Waves per SIMD the tile kernels are built for.  512-thread workgroups:
// two per CU for complex128 (64 KiB tile + 16 KiB matrices; at most 128
// VGPRs), four for complex64 (at most 64 VGPRs; its 32 KiB of LDS would
// allow five); programs with 5-qubit ops get more registers (complex128
// unconstrained, complex64 two workgroups per CU): their 5-qubit passes
// need them.  256-thread workgroups: half the waves per workgroup.  The
// complex64 kernels without 5-qubit ops use about 70 registers, so five of
// their workgroups fit a CU as their LDS allows (a bound of five per SIMD
// on the compiler measured 0.3% slower).
template <typename R, int KMAX, int NT>
constexpr int tile_waves_per_eu()
{
     const int w512 = sizeof(R) == 8 ? (KMAX >= 5 ? 1 : 4) : (KMAX >= 5 ? 4 : 8);
     return NT >= 512 ? w512 : (w512 > 1 ? w512 / 2 : 1);
}

// ops[0..nops) on the tile in LDS, a workgroup barrier before each.  The
// first `staged` elements of the ops' matrices are in smat (back to back in
// op order).
template <typename R, int KMAX, int NT, int MODE>
__device__ __forceinline__ void tile_ops(VT<R>* tile, const VT<R>* smat, const int staged,
                                         const VT<R>* __restrict__ mats,
                                         const TileOp* __restrict__ ops, const int nops)
{
     using V = VT<R>;
     constexpr int MK = tile_mfma_min_k(MODE);
     int soff = 0;
     for (int p = 0; p < nops; ++p) {
          const TileOp o = load_uniform(ops + p);
          const int sz = 1 << (2 * o.k);
          const bool in_lds = soff + sz <= staged;
          __syncthreads();  // the previous op (or the tile load) is complete
          if constexpr (MK <= 5) {
               if (o.k >= MK) {
                    if (in_lds) {
                         const uint32_t mo = static_cast<uint32_t>(soff);
                         if (o.k == 3) wg_pass_mfma<3, R, NT, true>(tile, smat, mo, o.rpos);
                         else if (o.k == 4) wg_pass_mfma<4, R, NT, true>(tile, smat, mo, o.rpos);
                         else if constexpr (KMAX >= 5) wg_pass_mfma<5, R, NT, true>(tile, smat, mo, o.rpos);
                    }
                    else {
                         const V* const m = mats + o.mat_off;
                         if (o.k == 3) wg_pass_mfma<3, R, NT, false>(tile, m, 0, o.rpos);
                         else if (o.k == 4) wg_pass_mfma<4, R, NT, false>(tile, m, 0, o.rpos);
                         else if constexpr (KMAX >= 5) wg_pass_mfma<5, R, NT, false>(tile, m, 0, o.rpos);
                    }
                    soff += sz;
                    continue;
               }
          }
          soff += sz;
          const V* const m = mats + o.mat_off;
          switch (o.k) {
               case 1: wg_pass_gm<1, R, NT>(tile, m, o.rpos); break;
               case 2: wg_pass_gm<2, R, NT>(tile, m, o.rpos); break;
               case 3:
                    if constexpr (MK > 3) wg_pass_gm<3, R, NT>(tile, m, o.rpos);
                    break;
               case 4:
                    if constexpr (MK > 4) wg_pass_gm<4, R, NT>(tile, m, o.rpos);
                    break;
               default:
                    if constexpr (KMAX >= 5 && MK > 5) wg_pass_gm<5, R, NT>(tile, m, o.rpos);
                    break;
          }
     }
}

// One tile program on tile `tix` of the state at psi.
template <typename R, int KMAX, int NT, int MODE>
__device__ __forceinline__ void tile_program(VT<R>* __restrict__ psi, const VT<R>* __restrict__ mats,
                                             const TileOp* __restrict__ ops, const int nops,
                                             const uint8_t (&hpos)[kTileHigh], const uint64_t tix,
                                             VT<R>* tile, VT<R>* smat)
{
     using V = VT<R>;
     using IO = TileIo<R, NT>;
     using MS = MatStage<R, NT, KMAX>;
     V v[IO::PER];
     IO::load(psi, tix, hpos, v);
     int staged = 0;
     constexpr bool kStage = MODE != kTileVector && MS::ME > 0;
     V mr[kStage ? MS::PER : 1];
     if constexpr (kStage) staged = MS::load(mr, mats, ops, nops);
     IO::to_lds(tile, v);
     if constexpr (kStage) MS::to_lds(smat, mr, staged);
     tile_ops<R, KMAX, NT, MODE>(tile, smat, staged, mats, ops, nops);
     __syncthreads();
     IO::from_lds(tile, v);
     IO::store(psi, tix, hpos, v);
}

// Tile pass over one state: workgroup b runs the program on tile
// swizzle(b).
template <typename R, int KMAX, int NT, int MODE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<R, KMAX, NT>())))
    k_tile_pass(VT<R>* __restrict__ psi, const VT<R>* __restrict__ mats,
                const TileOp* __restrict__ ops, const int nops, const TileHigh h, const int swz,
                const uint64_t nblocks)
{
     __shared__ VT<R> tile[1 << kTileBits];
     VT<R>* const smat = smat_buf<VT<R>, MODE != kTileVector ? tile_stage_elems<R, KMAX>() : 0>();
     const uint64_t b = block_linear();
     if (b >= nblocks) return;  // padding of the 2-D grid (whole block)
     tile_program<R, KMAX, NT, MODE>(psi, mats, ops, nops, h.hpos, swizzle_block(b, swz), tile, smat);
}

// Tile passes of many states: block b belongs to the item whose blocks
// contain it; the item's tiles are visited in swizzled order.
template <typename R, int KMAX, int NT, int MODE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<R, KMAX, NT>())))
    k_batch_tprog(VT<R>* __restrict__ arena, const VT<R>* __restrict__ mats,
                  const BatchTileProg* __restrict__ items, const int nitems,
                  const TileOp* __restrict__ ops, const uint64_t total_blocks, const int swz)
{
     __shared__ VT<R> tile[1 << kTileBits];
     VT<R>* const smat = smat_buf<VT<R>, MODE != kTileVector ? tile_stage_elems<R, KMAX>() : 0>();
     const uint64_t b = block_linear();
     if (b >= total_blocks) return;  // padding of the 2-D grid (whole block)
     const BatchTileProg it = load_uniform(items + find_item(items, nitems, b));
     const uint64_t tix = swizzle_block(b - it.block_start, swz < it.log_blocks ? swz : it.log_blocks);
     tile_program<R, KMAX, NT, MODE>(arena + it.state_off, mats, ops + it.op_first, it.nops, it.hpos, tix, tile,
                                   smat);
}

template <int K, typename R>
__global__ void __launch_bounds__(kBatchLayerBlock)
    k_batch_layer(VT<R>* __restrict__ arena, const VT<R>* __restrict__ mats,
                  const BatchItem* __restrict__ items, const int nitems, const uint64_t total_blocks,
                  const int swz)
{
     using V = VT<R>;
     constexpr int D = 1 << K;
     const uint64_t b = block_linear();
     if (b >= total_blocks) return;  // padding of the 2-D grid
     int lo = 0, hi = nitems - 1;
     while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (items[mid].block_start <= b) lo = mid;
          else hi = mid - 1;
     }
     const BatchItem it = items[lo];
     const int lbk = it.log_groups > 8 ? static_cast<int>(it.log_groups) - 8 : 0;
     const uint64_t t = swizzle_block(b - it.block_start, swz < lbk ? swz : lbk) * kBatchLayerBlock + threadIdx.x;
     if ((t >> it.log_groups) != 0) return;
     uint64_t base = t;
     uint64_t off[D];
#pragma unroll
     for (int l = 0; l < K; ++l) {
          const uint64_t lo_bits = base & ((1ull << it.pos[l]) - 1ull);
          base = ((base ^ lo_bits) << 1) | lo_bits;
     }
#pragma unroll
     for (int j = 0; j < D; ++j) {
          uint64_t o = 0;
#pragma unroll
          for (int l = 0; l < K; ++l)
               if ((j >> l) & 1) o |= 1ull << it.pos[l];
          off[j] = o;
     }
     V* const p = arena + it.state_off + base;
     V v[D];
#pragma unroll
     for (int j = 0; j < D; ++j) v[j] = p[off[j]];
     group_matvec<K, V, R>(mats + it.mat_off, v, [&](int i, V o) { p[off[i]] = o; });
}

template <int K, typename R>
__device__ __forceinline__ void layer_group(VT<R>* __restrict__ arena, const VT<R>* __restrict__ mats,
                                            const BatchItem& it, uint64_t t)
{
     using V = VT<R>;
     constexpr int D = 1 << K;
     uint64_t base = t;
     uint64_t off[D];
#pragma unroll
     for (int l = 0; l < K; ++l) {
          const uint64_t lo_bits = base & ((1ull << it.pos[l]) - 1ull);
          base = ((base ^ lo_bits) << 1) | lo_bits;
     }
#pragma unroll
     for (int j = 0; j < D; ++j) {
          uint64_t o = 0;
#pragma unroll
          for (int l = 0; l < K; ++l)
               if ((j >> l) & 1) o |= 1ull << it.pos[l];
          off[j] = o;
     }
     V* const p = arena + it.state_off + base;
     V v[D];
#pragma unroll
     for (int j = 0; j < D; ++j) v[j] = p[off[j]];
     group_matvec<K, V, R>(mats + it.mat_off, v, [&](int i, V o) { p[off[i]] = o; });
}

template <typename R, int KMAX>
__global__ void __launch_bounds__(kBatchLayerBlock)
    k_serve_layer(VT<R>* __restrict__ arena, const VT<R>* __restrict__ mats,
                  const BatchItem* __restrict__ items, const int nitems, const uint64_t total_blocks,
                  const int swz)
{
     const uint64_t b = block_linear();
     if (b >= total_blocks) return;  // padding of the 2-D grid
     const BatchItem it = items[find_item(items, nitems, b)];
     const int lbk = it.log_groups > 8 ? static_cast<int>(it.log_groups) - 8 : 0;
     const uint64_t t = swizzle_block(b - it.block_start, swz < lbk ? swz : lbk) * kBatchLayerBlock + threadIdx.x;
     if ((t >> it.log_groups) != 0) return;
     switch (it.k) {
          case 1: layer_group<1, R>(arena, mats, it, t); break;
          case 2: layer_group<2, R>(arena, mats, it, t); break;
          case 3: layer_group<3, R>(arena, mats, it, t); break;
          case 4: layer_group<4, R>(arena, mats, it, t); break;
          default:
               if constexpr (KMAX >= 5) layer_group<5, R>(arena, mats, it, t);
               break;
     }
}

// Batched wave-tile pass (the single-state k_dense_tile applied to the item
// the block belongs to).
template <int K, typename R, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    k_batch_tile(VT<R>* __restrict__ arena, const VT<R>* __restrict__ mats,
                 const BatchTileItem* __restrict__ items, const int nitems, const uint64_t total_blocks,
                 const int swz)
{
     using V = VT<R>;
     constexpr int D = 1 << K;
     __shared__ V tile[WPB][64 * D];
     const uint64_t b = block_linear();
     if (b >= total_blocks) return;  // padding of the 2-D grid (whole block)
     const BatchTileItem& it = items[find_item(items, nitems, b)];
     const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
     V* const T = tile[wv];
     const uint64_t lb = swizzle_block(b - it.block_start, swz < it.log_blocks ? swz : it.log_blocks);
     uint64_t idx = (lb * WPB + wv) << 6;
#pragma unroll
     for (int l = 0; l < K; ++l) {
          const uint64_t lo_bits = idx & ((1ull << it.ins[l]) - 1ull);
          idx = ((idx ^ lo_bits) << 1) | lo_bits;
     }
     V* const base = arena + it.state_off + idx + lane;
     V v[D];
#pragma unroll
     for (int i = 0; i < D; ++i) v[i] = ld<true>(base + it.hoff[i]);
#pragma unroll
     for (int i = 0; i < D; ++i) T[tile_slot<V>(i * 64 + lane)] = v[i];
     __syncthreads();
     int rl = 0;
#pragma unroll
     for (int q = 0; q < 6; ++q) rl |= ((lane >> q) & 1) << it.rank_n[q];
     int slot[D];
#pragma unroll
     for (int j = 0; j < D; ++j) {
          int r = rl;
#pragma unroll
          for (int l = 0; l < K; ++l)
               if ((j >> l) & 1) r |= 1 << it.rank_t[l];
          slot[j] = tile_slot<V>(r);
     }
#pragma unroll
     for (int j = 0; j < D; ++j) v[j] = T[slot[j]];
     __syncthreads();
     group_matvec<K, V, R>(mats + it.mat_off, v, [&](int i, V o) { T[slot[i]] = o; });
     __syncthreads();
#pragma unroll
     for (int i = 0; i < D; ++i) st<true>(base + it.hoff[i], T[tile_slot<V>(i * 64 + lane)]);
}

template <typename R>
__global__ void __launch_bounds__(kBatchLayerBlock)
    k_init_states(VT<R>* __restrict__ arena, const BatchItem* __restrict__ items, const int nitems,
                  const uint64_t total_blocks)
{
     using V = VT<R>;
     constexpr int E = kInitAmpsPerBlock / kBatchLayerBlock;
     const uint64_t b = block_linear();
     if (b >= total_blocks) return;  // padding of the 2-D grid
     const BatchItem it = items[find_item(items, nitems, b)];
     const uint64_t first = (b - it.block_start) * kInitAmpsPerBlock;
     const uint64_t N = 1ull << it.log_groups;
#pragma unroll
     for (int e = 0; e < E; ++e) {
          const uint64_t i = first + static_cast<uint64_t>(e) * kBatchLayerBlock + threadIdx.x;
          if (i < N) arena[it.state_off + i] = basis_amp<V>(i == 0);
     }
}

template <typename R>
__global__ void __launch_bounds__(kBatchLayerBlock)
    k_block_sums_multi(const VT<R>* __restrict__ arena, const BatchItem* __restrict__ items,
                       const int nitems, const uint64_t total_blocks, double* __restrict__ out)
{
     const uint64_t b = block_linear();
     if (b >= total_blocks) return;  // padding of the 2-D grid (whole block)
     const BatchItem it = items[find_item(items, nitems, b)];
     const uint64_t j = b - it.block_start;
     const uint64_t K = 1ull << it.log_groups;
     const VT<R>* const p = arena + it.state_off + j * K;
     double acc = 0.0;
     for (uint64_t i = threadIdx.x; i < K; i += kBatchLayerBlock) {
          const double re = p[i].x, im = p[i].y;
          acc += re * re + im * im;
     }
     __shared__ double red[kBatchLayerBlock / 64];
     for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
     if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
     __syncthreads();
     if (threadIdx.x == 0) {
          double r = 0;
          for (int w = 0; w < kBatchLayerBlock / 64; ++w) r += red[w];
          out[it.mat_off + j] = r;
     }
}

constexpr int kSampleThreads = 1024;

// One workgroup per job: running sum of the block probabilities (in place),
// then per shot u = U[0,1) * total, the first block whose running sum
// exceeds u, and a scan of that block's amplitudes.
template <typename R>
__global__ void __launch_bounds__(kSampleThreads)
    k_sample(const VT<R>* __restrict__ arena, const SampleJob* __restrict__ jobs,
             double* __restrict__ sums, uint64_t* __restrict__ out)
{
     const SampleJob jb = jobs[blockIdx.x];
     double* const cdf = sums + jb.sums_off;
     const uint32_t nb = 1u << jb.log_blocks;
     const uint32_t per = nb > kSampleThreads ? nb / kSampleThreads : 1u;
     const uint32_t t = threadIdx.x;
     const uint32_t first = t * per;
     double local = 0.0;
     if (first < nb)
          for (uint32_t i = 0; i < per; ++i) local += cdf[first + i];
     __shared__ double part[kSampleThreads];
     part[t] = local;
     __syncthreads();
     for (uint32_t o = 1; o < kSampleThreads; o <<= 1) {
          const double add = t >= o ? part[t - o] : 0.0;
          __syncthreads();
          part[t] += add;
          __syncthreads();
     }
     const double total = part[kSampleThreads - 1];
     if (first < nb) {
          double run = t ? part[t - 1] : 0.0;
          for (uint32_t i = 0; i < per; ++i) {
               run += cdf[first + i];
               cdf[first + i] = run;
          }
     }
     __syncthreads();
     const uint64_t K = 1ull << jb.log_block_amps;
     const VT<R>* const psi = arena + jb.state_off;
     const uint2 key = make_uint2(static_cast<uint32_t>(jb.seed), static_cast<uint32_t>(jb.seed >> 32));
     for (uint32_t s = t; s < jb.shots; s += kSampleThreads) {
          const uint4 r = philox4x32_10(make_uint4(s, 0u, 0x5a3du, 0u), key);
          const double u = uniform53(r.x, r.y) * total;
          uint32_t lo = 0, hi = nb - 1;  // first block with cdf > u
          while (lo < hi) {
               const uint32_t mid = (lo + hi) >> 1;
               if (cdf[mid] > u) hi = mid;
               else lo = mid + 1;
          }
          double rem = u - (lo ? cdf[lo - 1] : 0.0);
          const uint64_t b0 = static_cast<uint64_t>(lo) * K;
          uint64_t pick = b0, last_nz = b0;
          bool found = false;
          for (uint64_t i = 0; i < K; ++i) {
               const double re = psi[b0 + i].x, im = psi[b0 + i].y;
               const double pr = re * re + im * im;
               if (pr > 0.0) last_nz = b0 + i;
               if (rem < pr) {
                    pick = b0 + i;
                    found = true;
                    break;
               }
               rem -= pr;
          }
          out[jb.out_off + s] = found ? pick : last_nz;
     }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_set_ones(VT<R>* __restrict__ arena, const uint64_t* __restrict__ offsets, const int ncirc)
{
     const int c = blockIdx.x * kBlock + threadIdx.x;
     if (c >= ncirc) return;
     arena[offsets[c]] = basis_amp<VT<R>>(true);
}

}  // namespace

hipError_t launch_batch_tprog(void* arena, Prec prec, const void* mats, const BatchTileProg* items,
                              int nitems, const TileOp* ops, int kmax, uint64_t total_blocks,
                              hipStream_t s)
{
     if (nitems <= 0 || total_blocks == 0) return hipSuccess;
     if (kmax < 1 || kmax > 5) return hipErrorInvalidValue;
     const dim3 grid = grid_2d(total_blocks);
     const int swz = get_block_swizzle();
     const int mode = get_tile_mfma();
     const bool narrow = tile_threads_for(prec, kmax) == 256;
#define HQ_BTP(RR, KM, NT, MODE)                                                                   \
     hipLaunchKernelGGL((k_batch_tprog<RR, KM, NT, MODE>), grid, dim3(NT), 0, s,                   \
                        static_cast<VT<RR>*>(arena), static_cast<const VT<RR>*>(mats), items,     \
                        nitems, ops, total_blocks, swz)
#define HQ_BTP2(RR, KM)                                                \
     if (mode == kTileMatrix && narrow) HQ_BTP(RR, KM, 256, kTileMatrix); \
     else if (mode == kTileMatrix) HQ_BTP(RR, KM, 512, kTileMatrix);      \
     else HQ_BTP(RR, KM, 512, kTileVector)
     if (prec == Prec::C128) {
          if (kmax >= 5) HQ_BTP2(double, 5);
          else HQ_BTP2(double, 4);
     }
     else {
          if (kmax >= 5) HQ_BTP2(float, 5);
          else HQ_BTP2(float, 4);
     }
#undef HQ_BTP2
#undef HQ_BTP
     return hipGetLastError();
}

hipError_t launch_tile_pass(void* psi, Prec prec, int n_qubits, const TileHigh& h, const TileOp* ops,
                            int nops, int kmax, const void* mats, hipStream_t s)
{
     if (n_qubits < kTileBits || n_qubits > 40 || nops < 0 || kmax < 1 || kmax > 5)
          return hipErrorInvalidValue;
     for (int j = 0; j < kTileHigh; ++j)
          if (h.hpos[j] < 6 || h.hpos[j] >= n_qubits || (j && h.hpos[j] <= h.hpos[j - 1]))
               return hipErrorInvalidValue;
     const uint64_t blocks = 1ull << (n_qubits - kTileBits);
     const int bits = get_block_swizzle();
     const int swz = (bits > 0 && bits <= 16 && blocks % (1ull << bits) == 0) ? bits : 0;
     const dim3 grid = grid_2d(blocks);
     const int mode = get_tile_mfma();
     const bool narrow = tile_threads_for(prec, kmax) == 256;
#define HQ_TP(RR, KM, NT, MODE)                                                                    \
     hipLaunchKernelGGL((k_tile_pass<RR, KM, NT, MODE>), grid, dim3(NT), 0, s,                     \
                        static_cast<VT<RR>*>(psi), static_cast<const VT<RR>*>(mats), ops, nops, h, \
                        swz, blocks)
#define HQ_TP2(RR, KM)                                                \
     if (mode == kTileMatrix && narrow) HQ_TP(RR, KM, 256, kTileMatrix); \
     else if (mode == kTileMatrix) HQ_TP(RR, KM, 512, kTileMatrix);      \
     else HQ_TP(RR, KM, 512, kTileVector)
     if (prec == Prec::C128) {
          if (kmax >= 5) HQ_TP2(double, 5);
          else HQ_TP2(double, 4);
     }
     else {
          if (kmax >= 5) HQ_TP2(float, 5);
          else HQ_TP2(float, 4);
     }
#undef HQ_TP2
#undef HQ_TP
     return hipGetLastError();
}

int batch_wg_threads(Prec prec, int kmax) { return prec == Prec::C64 && kmax <= 4 ? 1024 : 512; }

hipError_t launch_batch_wg(void* arena, Prec prec, const void* mats, const BatchPass* passes,
                           const BatchCirc* circs, const uint32_t* wg_off,
                           const uint32_t* wg_list, int n_wg, int kmax, hipStream_t s)
{
     if (n_wg <= 0) return hipSuccess;
     if (kmax < 1 || kmax > 5) return hipErrorInvalidValue;
#define HQ_WG(RR, KM, NT)                                                                        \
     hipLaunchKernelGGL((k_batch_wg<RR, KM, NT>), dim3(n_wg), dim3(NT), 0, s,                     \
                        static_cast<VT<RR>*>(arena), static_cast<const VT<RR>*>(mats), passes,    \
                        circs, wg_off, wg_list)
     if (prec == Prec::C128) {
          if (kmax >= 5) HQ_WG(double, 5, 512);
          else HQ_WG(double, 4, 512);
     }
     else {
          if (kmax >= 5) HQ_WG(float, 5, 512);
          else HQ_WG(float, 4, 1024);
     }
#undef HQ_WG
     return hipGetLastError();
}

hipError_t launch_batch_layer(void* arena, Prec prec, const void* mats, const BatchItem* items,
                              int nitems, int k, uint64_t total_blocks, hipStream_t s)
{
     if (nitems <= 0) return hipSuccess;
     const dim3 grid = grid_2d(total_blocks);
#define HQ_LAYER(KK)                                                                             \
     do {                                                                                         \
          if (prec == Prec::C128)                                                                 \
               hipLaunchKernelGGL((k_batch_layer<KK, double>), grid, dim3(kBatchLayerBlock), 0,   \
                                  s, static_cast<double2*>(arena),                                \
                                  static_cast<const double2*>(mats), items, nitems, total_blocks, \
                                  get_block_swizzle_for(KK));                                     \
          else                                                                                    \
               hipLaunchKernelGGL((k_batch_layer<KK, float>), grid, dim3(kBatchLayerBlock), 0, s, \
                                  static_cast<float2*>(arena), static_cast<const float2*>(mats),  \
                                  items, nitems, total_blocks, get_block_swizzle_for(KK));        \
     } while (0)
     switch (k) {
          case 1: HQ_LAYER(1); break;
          case 2: HQ_LAYER(2); break;
          case 3: HQ_LAYER(3); break;
          case 4: HQ_LAYER(4); break;
          case 5: HQ_LAYER(5); break;
          default: return hipErrorInvalidValue;
     }
#undef HQ_LAYER
     return hipGetLastError();
}

hipError_t launch_serve_layer(void* arena, Prec prec, const void* mats, const BatchItem* items,
                              int nitems, uint64_t total_blocks, int kmax, hipStream_t s)
{
     if (nitems <= 0) return hipSuccess;
     if (kmax < 1 || kmax > 5) return hipErrorInvalidValue;
     const dim3 grid = grid_2d(total_blocks);
#define HQ_SERVE(RR, KM)                                                                         \
     hipLaunchKernelGGL((k_serve_layer<RR, KM>), grid, dim3(kBatchLayerBlock), 0, s,              \
                        static_cast<VT<RR>*>(arena), static_cast<const VT<RR>*>(mats), items,     \
                        nitems, total_blocks, get_block_swizzle())
     if (prec == Prec::C128) {
          if (kmax >= 5) HQ_SERVE(double, 5);
          else HQ_SERVE(double, 4);
     }
     else {
          if (kmax >= 5) HQ_SERVE(float, 5);
          else HQ_SERVE(float, 4);
     }
#undef HQ_SERVE
     return hipGetLastError();
}


hipError_t launch_batch_tile(void* arena, Prec prec, const void* mats, const BatchTileItem* items,
                             int nitems, int k, uint64_t total_blocks, hipStream_t s)
{
     if (nitems <= 0) return hipSuccess;
     const dim3 grid = grid_2d(total_blocks);
#define HQ_TILE(KK)                                                                              \
     do {                                                                                         \
          if (prec == Prec::C128) {                                                               \
               constexpr int W = (64 << KK) * 16 > 16384 ? 2 : 4;                                 \
               hipLaunchKernelGGL((k_batch_tile<KK, double, W>), grid, dim3(64 * W), 0, s,        \
                                  static_cast<double2*>(arena), static_cast<const double2*>(mats), \
                                  items, nitems, total_blocks, get_block_swizzle_for(KK));        \
          }                                                                                       \
          else {                                                                                  \
               constexpr int W = (64 << KK) * 8 > 16384 ? 2 : 4;                                  \
               hipLaunchKernelGGL((k_batch_tile<KK, float, W>), grid, dim3(64 * W), 0, s,         \
                                  static_cast<float2*>(arena), static_cast<const float2*>(mats),   \
                                  items, nitems, total_blocks, get_block_swizzle_for(KK));        \
          }                                                                                       \
     } while (0)
     switch (k) {
          case 1: HQ_TILE(1); break;
          case 2: HQ_TILE(2); break;
          case 3: HQ_TILE(3); break;
          case 4: HQ_TILE(4); break;
          case 5: HQ_TILE(5); break;
          default: return hipErrorInvalidValue;
     }
#undef HQ_TILE
     return hipGetLastError();
}

hipError_t launch_init_states(void* arena, Prec prec, const BatchItem* items, int nitems,
                              uint64_t total_blocks, hipStream_t s)
{
     if (nitems <= 0) return hipSuccess;
     const dim3 grid = grid_2d(total_blocks);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_init_states<double>), grid, dim3(kBatchLayerBlock), 0, s,
                             static_cast<double2*>(arena), items, nitems, total_blocks);
     else
          hipLaunchKernelGGL((k_init_states<float>), grid, dim3(kBatchLayerBlock), 0, s,
                             static_cast<float2*>(arena), items, nitems, total_blocks);
     return hipGetLastError();
}

hipError_t launch_block_sums_multi(const void* arena, Prec prec, const BatchItem* items,
                                   int nitems, uint64_t total_blocks, double* out, hipStream_t s)
{
     if (nitems <= 0) return hipSuccess;
     const dim3 grid = grid_2d(total_blocks);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_block_sums_multi<double>), grid, dim3(kBatchLayerBlock), 0, s,
                             static_cast<const double2*>(arena), items, nitems, total_blocks, out);
     else
          hipLaunchKernelGGL((k_block_sums_multi<float>), grid, dim3(kBatchLayerBlock), 0, s,
                             static_cast<const float2*>(arena), items, nitems, total_blocks, out);
     return hipGetLastError();
}

hipError_t launch_sample(const void* arena, Prec prec, const SampleJob* jobs, int njobs,
                         double* sums, uint64_t* out, hipStream_t s)
{
     if (njobs <= 0) return hipSuccess;
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_sample<double>), dim3(njobs), dim3(kSampleThreads), 0, s,
                             static_cast<const double2*>(arena), jobs, sums, out);
     else
          hipLaunchKernelGGL((k_sample<float>), dim3(njobs), dim3(kSampleThreads), 0, s,
                             static_cast<const float2*>(arena), jobs, sums, out);
     return hipGetLastError();
}

hipError_t launch_set_ones(void* arena, Prec prec, const uint64_t* offsets, int ncirc, hipStream_t s)
{
     if (ncirc <= 0) return hipSuccess;
     const unsigned grid = static_cast<unsigned>((ncirc + kBlock - 1) / kBlock);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_set_ones<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(arena), offsets, ncirc);
     else
          hipLaunchKernelGGL((k_set_ones<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(arena), offsets, ncirc);
     return hipGetLastError();
}

}  // namespace hq
