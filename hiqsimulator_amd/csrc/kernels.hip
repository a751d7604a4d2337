// gfx950 (CDNA4) kernels of the state-vector hot path.
//
// Every kernel is one streaming pass over the state in HBM; the bound is HBM
// bandwidth (algorithmic bytes: one read and one write of each amplitude a
// pass touches).  The fused k-qubit pass has two forms:
//  * wave-tile (default): a wave owns 64 groups (the 2^k amplitudes one
//    fused matrix mixes, one group per lane).  Their 64 * 2^k amplitudes are
//    read as 2^k fully contiguous 64-amplitude chunks (1 KiB for complex128)
//    whatever the target bits are, regrouped through LDS, multiplied in
//    registers, regrouped back and written as the same chunks.  Global
//    accesses are non-temporal (the state is streamed once per pass).
//  * direct: each lane loads its own group straight from HBM; used when the
//    target bits are all high (the chunks of a tile then lie far apart and
//    the LDS round trip costs more than it saves) or when a low control bit
//    breaks the chunk layout.
// The fused matrix is uniform across the grid and is read with scalar loads
// into SGPRs, which v_fma_f64 / v_fma_f32 take directly as an operand.
#include <hip/hip_runtime.h>

#include <atomic>

#include <cstdint>

#include "device_common.hpp"
#include "kernels.hpp"

namespace hq {
namespace {

// ---------------------------------------------------------------------------
// Dense fused k-qubit pass, direct form.  Restates kernels/intrin/
// kernel{1..5}.hpp (kernel_core + kernel_compute): for each base I with
// target bits 0 and control bits 1, v_j = psi[I + off_j],
// psi[I + off_i] = sum_j m[i][j] v_j.  GPT groups per lane are loaded before
// any is computed so that GPT * 2^K loads are in flight per lane.
// ---------------------------------------------------------------------------
template <int K, typename R, int GPT, bool NT, int BS = kBlock>
__global__ void __launch_bounds__(BS)
    k_dense_direct(typename C2<R>::T* __restrict__ psi, const typename C2<R>::T* __restrict__ mat,
                   const DenseDesc d, const uint64_t ngroups, const int swz)
{
     using V = typename C2<R>::T;
     constexpr int D = 1 << K;
     uint64_t off[D];
#pragma unroll
     for (int j = 0; j < D; ++j) {
          uint64_t o = 0;
#pragma unroll
          for (int l = 0; l < K; ++l)
               if ((j >> l) & 1) o += 1ull << d.pos[l];
          off[j] = o;
     }
     const uint64_t t0 = swizzle_block(block_linear(), swz) * (BS * GPT) + threadIdx.x;
     uint64_t I[GPT];
     V v[GPT][D];
#pragma unroll
     for (int g = 0; g < GPT; ++g) {
          const uint64_t t = t0 + static_cast<uint64_t>(g) * BS;
          I[g] = insert_zeros(t < ngroups ? t : 0, d.ins) | d.ctrlmask;
          if (t < ngroups) {
#pragma unroll
               for (int j = 0; j < D; ++j) v[g][j] = ld<NT>(psi + I[g] + off[j]);
          }
     }
#pragma unroll
     for (int g = 0; g < GPT; ++g) {
          if (t0 + static_cast<uint64_t>(g) * BS >= ngroups) continue;
          V* const p = psi + I[g];
          group_matvec<K, V, R>(mat, v[g], [&](int i, V o) { st<NT>(p + off[i], o); });
     }
}

// Wave-tile description (built on the host).  The tile index r of an
// amplitude is its index bits at the tile positions V = (targets) + (the six
// lowest non-inserted positions), in ascending bit order; V always holds
// bits 0..5 here, so amplitude r of a wave's tile sits at
// base + (r & 63) + hoff[r >> 6].
struct TileDesc {
     uint64_t hoff[32];   // address offset of chunk i = pdep(i, V above bit 5)
     uint8_t rank_n[6];   // tile rank of the group index's six lowest bits
     uint8_t rank_t[5];   // tile rank of matrix bit l
};

template <int K, typename R, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    k_dense_tile(typename C2<R>::T* __restrict__ psi, const typename C2<R>::T* __restrict__ mat,
                 const DenseDesc d, const TileDesc L, const int swz)
{
     using V = typename C2<R>::T;
     constexpr int D = 1 << K;
     __shared__ V tile[WPB][64 * D];
     const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
     V* const T = tile[wv];
     const uint64_t wave = swizzle_block(block_linear(), swz) * WPB + wv;
     V* const base = psi + (insert_zeros(wave << 6, d.ins) | d.ctrlmask) + lane;
     V v[D];
#pragma unroll
     for (int i = 0; i < D; ++i) v[i] = ld<true>(base + L.hoff[i]);
#pragma unroll
     for (int i = 0; i < D; ++i) T[tile_slot<V>(i * 64 + lane)] = v[i];
     __syncthreads();
     int rl = 0;
#pragma unroll
     for (int b = 0; b < 6; ++b) rl |= ((lane >> b) & 1) << L.rank_n[b];
     int slot[D];
#pragma unroll
     for (int j = 0; j < D; ++j) {
          int r = rl;
#pragma unroll
          for (int l = 0; l < K; ++l)
               if ((j >> l) & 1) r |= 1 << L.rank_t[l];
          slot[j] = tile_slot<V>(r);
     }
#pragma unroll
     for (int j = 0; j < D; ++j) v[j] = T[slot[j]];
     __syncthreads();
     group_matvec<K, V, R>(mat, v, [&](int i, V o) { T[slot[i]] = o; });
     __syncthreads();
#pragma unroll
     for (int i = 0; i < D; ++i) st<true>(base + L.hoff[i], T[tile_slot<V>(i * 64 + lane)]);
}

template <typename R>
struct DiagVals {
     R v[64];  // 32 complex values, interleaved
};

// Diagonal fused pass: kernels/intrin/kernels_diag.hpp:34-144.  Threads
// enumerate the indices whose control bits are 1 (zero bits inserted at the
// control positions, ctrlmask OR-ed in); d(i) gathers the target bits.
template <int K, typename R, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_diag(typename C2<R>::T* __restrict__ psi, const DiagVals<R> dv, const InsertSet cins,
           const uint64_t ctrlmask, const uint32_t pos0, const uint32_t pos1, const uint32_t pos2,
           const uint32_t pos3, const uint32_t pos4, const uint64_t count)
{
     using V = typename C2<R>::T;
     constexpr int E = 4;
     const uint32_t pos[5] = {pos0, pos1, pos2, pos3, pos4};
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * (kBlock * E);
     for (uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * (kBlock * E) + threadIdx.x; t0 < count;
          t0 += stride) {
#pragma unroll
          for (int e = 0; e < E; ++e) {
               const uint64_t t = t0 + static_cast<uint64_t>(e) * kBlock;
               if (t >= count) break;
               const uint64_t i = insert_zeros(t, cins) | ctrlmask;
               uint32_t di = 0;
#pragma unroll
               for (int l = 0; l < K; ++l) di |= static_cast<uint32_t>((i >> pos[l]) & 1ull) << l;
               const V a = ld<NT>(psi + i);
               const R mr = dv.v[2 * di], mi = dv.v[2 * di + 1];
               V o;
               o.x = fma(a.x, mr, -a.y * mi);
               o.y = fma(a.x, mi, a.y * mr);
               st<NT>(psi + i, o);
          }
     }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_scale(typename C2<R>::T* __restrict__ psi, const R fr, const R fi, const uint64_t n)
{
     using V = typename C2<R>::T;
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
          const V a = ld<true>(psi + i);
          V o;
          o.x = fma(a.x, fr, -a.y * fi);
          o.y = fma(a.x, fi, a.y * fr);
          st<true>(psi + i, o);
     }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_fill(typename C2<R>::T* __restrict__ psi, const R vr, const R vi, const uint64_t n)
{
     using V = typename C2<R>::T;
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     V o;
     o.x = vr;
     o.y = vi;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
          st<true>(psi + i, o);
}

// Plain streaming copy in 16-byte words.
__global__ void __launch_bounds__(kBlock)
    k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, const uint64_t nwords)
{
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < nwords;
          i += stride)
          dst[i] = src[i];
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_normalize(typename C2<R>::T* __restrict__ psi, const uint64_t n, const uint64_t msk,
                const uint64_t val, const R f)
{
     using V = typename C2<R>::T;
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
          V a = ld<true>(psi + i);
          if ((i & msk) != val) {
               a.x = 0;
               a.y = 0;
          }
          else {
               a.x *= f;
               a.y *= f;
          }
          st<true>(psi + i, a);
     }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_remove_bit(typename C2<R>::T* __restrict__ dst, const typename C2<R>::T* __restrict__ src,
                 const uint64_t n_dst, const int pos, const uint64_t keep)
{
     const uint64_t lo_mask = (1ull << pos) - 1ull;
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < n_dst;
          j += stride) {
          const uint64_t lo = j & lo_mask;
          const uint64_t i = ((j ^ lo) << 1) | lo | (keep << pos);
          dst[j] = src[i];
     }
}

// Qubit-index permutation: dst[j] = src[i(j)], bit p of j = bit src_pos[p] of
// i.  i(j) is assembled from five 8-bit lookup tables built in LDS.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_permute(typename C2<R>::T* __restrict__ dst, const typename C2<R>::T* __restrict__ src,
              const uint64_t n, const int nq, const InsertSet sp)
{
     __shared__ uint64_t tab[5][256];
     for (int c = 0; c < 5; ++c) {
          const int v = threadIdx.x;
          uint64_t r = 0;
          for (int b = 0; b < 8; ++b) {
               const int p = 8 * c + b;
               if (p < nq && ((v >> b) & 1)) r |= 1ull << sp.p[p];
          }
          tab[c][v] = r;
     }
     __syncthreads();
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < n; j += stride) {
          const uint64_t i = tab[0][j & 255] | tab[1][(j >> 8) & 255] | tab[2][(j >> 16) & 255]
                             | tab[3][(j >> 24) & 255] | tab[4][(j >> 32) & 255];
          dst[j] = src[i];
     }
}

// Tiled qubit permutation (LDS transpose).  The tile bits U (dst index
// positions, ascending) contain dst bits 0..5 and the dst positions of src
// bits 0..5, so both sides of the tile are read/written as contiguous
// 64-amplitude runs.  Tile element r (src order: bit a of r = src bit
// src_sorted[a]) lands at tile slot q (dst order: bit a of q = dst bit
// u_sorted[a]); q is r with its bits moved by qbit_of_rbit.
struct PermTile {
     InsertSet u;           // tile dst positions, ascending (= insert set)
     uint8_t src_sorted[12];  // src positions of the tile bits, ascending
     uint8_t qbit_of_rbit[12];
     uint8_t src_pos[40];   // dst bit p <- src bit src_pos[p]
     int nq;
};

template <typename V>
__device__ __forceinline__ int perm_slot(int q)
{
     return q ^ ((q >> 4) & 15) ^ ((q >> 8) & 15);
}

// Workgroups of 2^TB threads; tile `tix` = the block index with its low
// `swz` bits reversed (swizzle_block), so the blocks in flight spread over
// far-apart source addresses.
template <typename R, int T, int TB>
__global__ void __launch_bounds__(1 << TB)
    k_permute_tile(typename C2<R>::T* __restrict__ dst, const typename C2<R>::T* __restrict__ src,
                   const PermTile P, const int swz)
{
     using V = typename C2<R>::T;
     constexpr int NT = 1 << TB;
     constexpr int E = 1 << (T - TB);
     __shared__ V tile[1 << T];
     const uint64_t dbase = insert_zeros(swizzle_block(block_linear(), swz), P.u);
     uint64_t sbase = 0;
     for (int p = 0; p < P.nq; ++p)
          if ((dbase >> p) & 1) sbase |= 1ull << P.src_pos[p];
     const int tid = threadIdx.x;
     uint64_t slo = 0, dlo = 0;
     int qlo = 0;
#pragma unroll
     for (int b = 0; b < TB; ++b) {
          const uint64_t bit = (tid >> b) & 1;
          slo |= bit << P.src_sorted[b];
          dlo |= bit << P.u.p[b];
          qlo |= static_cast<int>(bit) << P.qbit_of_rbit[b];
     }
#pragma unroll
     for (int e = 0; e < E; ++e) {
          uint64_t shi = 0;
          int qhi = 0;
#pragma unroll
          for (int b = 0; b < T - TB; ++b) {
               const uint64_t bit = (e >> b) & 1;
               shi |= bit << P.src_sorted[TB + b];
               qhi |= static_cast<int>(bit) << P.qbit_of_rbit[TB + b];
          }
          tile[perm_slot<V>(qlo | qhi)] = ld<true>(src + sbase + slo + shi);
     }
     __syncthreads();
#pragma unroll
     for (int e = 0; e < E; ++e) {
          uint64_t dhi = 0;
#pragma unroll
          for (int b = 0; b < T - TB; ++b) dhi |= static_cast<uint64_t>((e >> b) & 1) << P.u.p[TB + b];
          st<true>(dst + dbase + dlo + dhi, tile[perm_slot<V>(e * NT + tid)]);
     }
}

// dst[j] += c * (-1)^{popcount((j ^ x) & z)} * src[j ^ x].
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_pauli_acc(typename C2<R>::T* __restrict__ dst, const typename C2<R>::T* __restrict__ src,
                const uint64_t n, const uint64_t x, const uint64_t z, const R cr, const R ci)
{
     using V = typename C2<R>::T;
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < n; j += stride) {
          const uint64_t i = j ^ x;
          const V a = src[i];
          const R s = (__popcll(i & z) & 1) ? R(-1) : R(1);
          V o = dst[j];
          o.x += s * (cr * a.x - ci * a.y);
          o.y += s * (cr * a.y + ci * a.x);
          dst[j] = o;
     }
}

// ---------------------------------------------------------------------------
// Reductions: fixed grid (<= kReduceBlocks blocks), per-thread FP64
// accumulation in a grid-stride loop, wave shuffle + LDS tree inside a block,
// one partial per block, then a single-block final tree.  Deterministic for a
// given size.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
     for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
     return v;
}

template <int NOUT>
__device__ __forceinline__ void block_store(double (&acc)[NOUT], double* partials, int nblocks)
{
     __shared__ double red[NOUT][kBlock / 64];
     const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
     for (int o = 0; o < NOUT; ++o) {
          const double s = wave_sum(acc[o]);
          if (lane == 0) red[o][w] = s;
     }
     __syncthreads();
     if (threadIdx.x == 0) {
#pragma unroll
          for (int o = 0; o < NOUT; ++o) {
               double s = 0;
               for (int i = 0; i < kBlock / 64; ++i) s += red[o][i];
               partials[o * nblocks + blockIdx.x] = s;
          }
     }
}

template <int NOUT>
__global__ void __launch_bounds__(kBlock)
    k_reduce_final(const double* __restrict__ partials, const int nblocks, double* __restrict__ out)
{
     double acc[NOUT];
#pragma unroll
     for (int o = 0; o < NOUT; ++o) {
          double s = 0;
          for (int i = threadIdx.x; i < nblocks; i += kBlock) s += partials[o * nblocks + i];
          acc[o] = s;
     }
     __shared__ double red[NOUT][kBlock / 64];
     const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
     for (int o = 0; o < NOUT; ++o) {
          const double s = wave_sum(acc[o]);
          if (lane == 0) red[o][w] = s;
     }
     __syncthreads();
     if (threadIdx.x == 0) {
#pragma unroll
          for (int o = 0; o < NOUT; ++o) {
               double s = 0;
               for (int i = 0; i < kBlock / 64; ++i) s += red[o][i];
               out[o] = s;
          }
     }
}

template <typename R>
__device__ __forceinline__ double norm2(const typename C2<R>::T a)
{
     const double x = a.x, y = a.y;
     return x * x + y * y;
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_masked_norm(const typename C2<R>::T* __restrict__ psi, const uint64_t count,
                  const InsertSet ins, const uint64_t val, double* __restrict__ partials)
{
     double acc[1] = {0.0};
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; t < count;
          t += stride)
          acc[0] += norm2<R>(psi[insert_zeros(t, ins) | val]);
     block_store<1>(acc, partials, gridDim.x);
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_two_bucket(const typename C2<R>::T* __restrict__ psi, const uint64_t n, const int pos,
                 double* __restrict__ partials)
{
     double acc[2] = {0.0, 0.0};
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
          const double p = norm2<R>(psi[i]);
          if ((i >> pos) & 1) acc[1] += p;
          else acc[0] += p;
     }
     block_store<2>(acc, partials, gridDim.x);
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_entropy(const typename C2<R>::T* __restrict__ psi, const uint64_t n,
              double* __restrict__ partials)
{
     double acc[1] = {0.0};
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
          const double a = hypot(static_cast<double>(psi[i].x), static_cast<double>(psi[i].y));
          const double pr = a * a;
          if (pr > 0) acc[0] += pr * log2(pr);
     }
     block_store<1>(acc, partials, gridDim.x);
}

// sum_i conj(psi[i ^ x]) psi[i] (-1)^{popcount(i & z)}; the uniform factor
// i^{popcount(x & z)} is applied by the host.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_pauli_exp(const typename C2<R>::T* __restrict__ psi, const uint64_t n, const uint64_t x,
                const uint64_t z, double* __restrict__ partials)
{
     double acc[2] = {0.0, 0.0};
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
          const double ar = psi[i].x, ai = psi[i].y;
          const double br = psi[i ^ x].x, bi = psi[i ^ x].y;
          const double s = (__popcll(i & z) & 1) ? -1.0 : 1.0;
          acc[0] += s * (br * ar + bi * ai);
          acc[1] += s * (br * ai - bi * ar);
     }
     block_store<2>(acc, partials, gridDim.x);
}

// psi[i] = (g0 + i g1) with g0, g1 ~ N(0, 1) from Box-Muller on the Philox
// block of counter (i, seed_hi); the caller normalises afterwards.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_random_gaussian(typename C2<R>::T* __restrict__ psi, const uint64_t n, const uint64_t seed)
{
     using V = typename C2<R>::T;
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     const uint2 key = make_uint2(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
          const uint4 r = philox4x32_10(
              make_uint4(static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), 0x5eedu, 0u), key);
          // 53-bit uniforms in (0, 1].
          const double u0 = ((static_cast<uint64_t>(r.x) << 21 ^ r.y) + 1.0) * 0x1p-53;
          const double u1 = ((static_cast<uint64_t>(r.z) << 21 ^ r.w) + 1.0) * 0x1p-53;
          const double rad = sqrt(-2.0 * log(u0));
          double sn, cs;
          sincospi(2.0 * u1, &sn, &cs);
          V o;
          o.x = static_cast<R>(rad * cs);
          o.y = static_cast<R>(rad * sn);
          st<true>(psi + i, o);
     }
}

// Comparison with a reference copy: out[0..1] = sum conj(ref[i]) psi[i],
// out[2] = sum |ref[i]|^2, out[3] = max |psi[i] - ref[i]| (FP64).
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_compare(const typename C2<R>::T* __restrict__ psi, const typename C2<R>::T* __restrict__ ref,
              const uint64_t n, double* __restrict__ partials)
{
     double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
     const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
     for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
          const double ar = psi[i].x, ai = psi[i].y, br = ref[i].x, bi = ref[i].y;
          acc[0] += br * ar + bi * ai;
          acc[1] += br * ai - bi * ar;
          acc[2] += br * br + bi * bi;
          acc[4] += ar * ar + ai * ai;
          const double dr = ar - br, di = ai - bi;
          acc[3] = fmax(acc[3], sqrt(dr * dr + di * di));
     }
     // sums for 0, 1, 2, 4; max for 3
     __shared__ double red[5][kBlock / 64];
     const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
     for (int o = 0; o < 5; ++o) {
          if (o == 3) continue;
          const double sum = wave_sum(acc[o]);
          if (lane == 0) red[o][w] = sum;
     }
     double m = acc[3];
     for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
     if (lane == 0) red[3][w] = m;
     __syncthreads();
     if (threadIdx.x == 0) {
          for (int o = 0; o < 5; ++o) {
               double r = 0;
               for (int i = 0; i < kBlock / 64; ++i) r = o == 3 ? fmax(r, red[o][i]) : r + red[o][i];
               partials[o * gridDim.x + blockIdx.x] = r;
          }
     }
}

__global__ void __launch_bounds__(kBlock)
    k_compare_final(const double* __restrict__ partials, const int nblocks, double* __restrict__ out)
{
     if (threadIdx.x != 0) return;
     double s[5] = {0, 0, 0, 0, 0};
     for (int i = 0; i < nblocks; ++i)
          for (int o = 0; o < 5; ++o)
               s[o] = o == 3 ? fmax(s[o], partials[o * nblocks + i]) : s[o] + partials[o * nblocks + i];
     for (int o = 0; o < 5; ++o) out[o] = s[o];
}

// Measurement block sums (SimulatorMPI.cpp:817-839): one workgroup per
// contiguous block when the block is long, one thread per block otherwise.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_block_sums_wg(const typename C2<R>::T* __restrict__ psi, const uint64_t K,
                    double* __restrict__ out)
{
     const uint64_t b = blockIdx.x;
     double acc[1] = {0.0};
     for (uint64_t j = threadIdx.x; j < K; j += kBlock) acc[0] += norm2<R>(psi[b * K + j]);
     __shared__ double red[kBlock / 64];
     const double s = wave_sum(acc[0]);
     if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
     __syncthreads();
     if (threadIdx.x == 0) {
          double r = 0;
          for (int i = 0; i < kBlock / 64; ++i) r += red[i];
          out[b] = r;
     }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_block_sums_seq(const typename C2<R>::T* __restrict__ psi, const uint64_t K, const uint64_t nb,
                     double* __restrict__ out)
{
     const uint64_t b = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
     if (b >= nb) return;
     double s = 0;
     for (uint64_t j = 0; j < K; ++j) s += norm2<R>(psi[b * K + j]);
     out[b] = s;
}

// Waves per block of the tile kernel: LDS holds 64 * 2^K amplitudes per wave.
template <int K, typename R>
constexpr int tile_wpb()
{
     return (64 << K) * static_cast<int>(sizeof(typename C2<R>::T)) > 16384 ? 2 : 4;
}

// Dense passes whose targets all sit at or above this bit take the direct
// form.  With the swizzled block order the tile form is as fast or faster
// at every placement measured on MI355X (30 qubits, k = 1..4), so the
// default never picks the direct form for that reason.
// The knobs are process-wide and may be set from any thread while other
// threads launch: each is an atomic read once per launch.
std::atomic<int> g_high_target{64};
// Low block-index bits reversed by swizzle_block: for passes of k <= 3 (and
// the other streaming kernels) and for passes of k >= 4.  Measured on a
// 30-qubit k = 4 pass over 30 placements: 5421 GB/s average without
// swizzle, 5821 with 8 bits, 5940 with 6 bits.
std::atomic<int> g_swizzle_bits{8};
std::atomic<int> g_swizzle_bits_wide{6};
std::atomic<int> g_tile_mfma{1};
std::atomic<int> g_dense5_tile{1};
// 4-qubit passes with every target at bit 13 or above take the direct form
// in 1024-thread blocks (16 KiB contiguous per target row): 6.34 against
// 6.63 ms in the tile form on a 30-qubit state, 32 such placements.
std::atomic<int> g_high_target_wide{13};
std::atomic<int> g_direct_block{1024};
// Tiled qubit permutation: low block-index bits reversed, and threads per
// workgroup (256 or 512).  Measured on 16 random 30-qubit permutations, the
// same sequence for each setting (3 rounds): complex128 6.27 ms without the
// swizzle, 6.08 ms with 8 bits (5477 against 5650 GB/s), 6.37 ms with 8
// bits and 512 threads; complex64 3.55 against 3.53 ms.  (A 13-bit tile in
// 1024-thread workgroups, 2 KiB contiguous reads, measured 6.11 ms.)
std::atomic<int> g_permute_swizzle{8};
std::atomic<int> g_permute_block{256};
// Threads per workgroup of the tile-program kernels on the matrix cores: 0
// picks per precision.  Measured on the 4096-circuit batch: complex64 359
// ms with 256 threads against 376 with 512 (the per-op lane setup is shared
// by four column blocks per wave instead of two); complex128 693 against
// 680 ms, and the 30-qubit tile pass 8.07 against 7.85 ms (two workgroups
// per CU leave too few waves at 256).  Complex128 programs with 5-qubit ops
// take 256: one-op 5-qubit passes 6.67 against 6.85 ms on 30 qubits (each
// wave then holds two 32-row column blocks that share its operand loads).
std::atomic<int> g_tile_threads{0};

// Swizzle width for a grid of `blocks` blocks of a k-qubit pass: the
// configured width when the grid holds whole runs of 2^bits blocks, else
// none.
int swizzle_bits_for(uint64_t blocks, int k)
{
     const int bits = get_block_swizzle_for(k);
     if (bits <= 0 || bits > 16) return 0;
     return (blocks % (1ull << bits)) == 0 ? bits : 0;
}

// Direct form in blocks of BS threads (each block covers BS consecutive
// groups, so BS * sizeof(amplitude) contiguous bytes per target row).
template <int K, typename R, int GPT, int BS>
hipError_t dense_direct_launch(void* psi, const void* mat, const DenseDesc& d, uint64_t ngroups, bool nt,
                               hipStream_t s)
{
     using V = typename C2<R>::T;
     const uint64_t dblocks = (ngroups + BS * GPT - 1) / (BS * GPT);
     const int swz = (ngroups % (BS * GPT)) == 0 ? swizzle_bits_for(dblocks, K) : 0;
     if (nt)
          hipLaunchKernelGGL((k_dense_direct<K, R, GPT, true, BS>), grid_2d(dblocks), dim3(BS), 0, s,
                             static_cast<V*>(psi), static_cast<const V*>(mat), d, ngroups, swz);
     else
          hipLaunchKernelGGL((k_dense_direct<K, R, GPT, false, BS>), grid_2d(dblocks), dim3(BS), 0, s,
                             static_cast<V*>(psi), static_cast<const V*>(mat), d, ngroups, swz);
     return hipGetLastError();
}

template <int K, typename R>
hipError_t dense_k(void* psi, int n, const void* mat, const DenseDesc& d, hipStream_t s)
{
     using V = VT<R>;
     const int nins = d.ins.n;
     if (nins > n) return hipErrorInvalidValue;
     const uint64_t ngroups = 1ull << (n - nins);
     uint64_t ins_mask = 0, tmask = 0;
     for (int i = 0; i < nins; ++i) ins_mask |= 1ull << d.ins.p[i];
     int min_t = 64;
     for (int l = 0; l < K; ++l) {
          tmask |= 1ull << d.pos[l];
          min_t = d.pos[l] < static_cast<uint32_t>(min_t) ? static_cast<int>(d.pos[l]) : min_t;
     }
     constexpr int WPB = tile_wpb<K, R>();
     // Tile form: bits 0..5 must all be targets or free (no control among
     // them) and there must be at least one full block of waves.
     const bool low_ctrl = ((ins_mask & ~tmask) & 63ull) != 0;
     const bool tileable = !low_ctrl && ngroups >= 64ull * WPB;
     // For k = 5 the tile holds 32 KiB (complex128) per wave, which caps the
     // occupancy; the direct form is faster there unless bit 0 (or, for
     // complex64, bit 0 or 1) is a target.
     const bool tile_pays = K < 5 || min_t <= (sizeof(V) == 16 ? 0 : 1);
     if (tileable && tile_pays && min_t < get_tile_high_target_for(K)) {
          TileDesc L{};
          uint64_t lanes = 0;
          int nb = 0;
          for (int p = 0; p < n && nb < 6; ++p)
               if (!((ins_mask >> p) & 1)) {
                    lanes |= 1ull << p;
                    ++nb;
               }
          const uint64_t V_bits = lanes | tmask;
          int vpos[64], nv = 0, rank_of[64];
          for (int p = 0; p < 64; ++p)
               if ((V_bits >> p) & 1) {
                    rank_of[p] = nv;
                    vpos[nv++] = p;
               }
          if (nv != 6 + K) return hipErrorInvalidValue;
          nb = 0;
          for (int p = 0; p < 64; ++p)
               if ((lanes >> p) & 1) L.rank_n[nb++] = static_cast<uint8_t>(rank_of[p]);
          for (int l = 0; l < K; ++l) L.rank_t[l] = static_cast<uint8_t>(rank_of[d.pos[l]]);
          for (int i = 0; i < (1 << K); ++i) {
               uint64_t o = 0;
               for (int b = 0; b < K; ++b)
                    if ((i >> b) & 1) o |= 1ull << vpos[6 + b];
               L.hoff[i] = o;
          }
          const uint64_t blocks = ngroups / (64ull * WPB);
          hipLaunchKernelGGL((k_dense_tile<K, R, WPB>), grid_2d(blocks), dim3(64 * WPB), 0, s,
                             static_cast<V*>(psi), static_cast<const V*>(mat), d, L,
                             swizzle_bits_for(blocks, K));
          return hipGetLastError();
     }
     // Direct form; non-temporal when every lane access covers whole 128-byte
     // lines (no inserted bit below 128 / sizeof(V) amplitudes).
     constexpr int GPT = K >= 4 ? 1 : (K == 3 ? 4 : (K == 2 ? 8 : 16));
     const int line_bits = sizeof(V) == 16 ? 3 : 4;
     const bool nt = (ins_mask & ((1ull << line_bits) - 1)) == 0;
     if constexpr (K == 4) {
          const int bs = get_direct_block();
          if (bs == 512) return dense_direct_launch<K, R, GPT, 512>(psi, mat, d, ngroups, nt, s);
          if (bs == 1024) return dense_direct_launch<K, R, GPT, 1024>(psi, mat, d, ngroups, nt, s);
     }
     const uint64_t dblocks = (ngroups + kBlock * GPT - 1) / (kBlock * GPT);
     const dim3 grid = grid_2d(dblocks);
     // whole blocks only (a partial last block must stay last)
     const int swz = (ngroups % (kBlock * GPT)) == 0 ? swizzle_bits_for(dblocks, K) : 0;
     if (nt)
          hipLaunchKernelGGL((k_dense_direct<K, R, GPT, true>), grid, dim3(kBlock), 0, s,
                             static_cast<V*>(psi), static_cast<const V*>(mat), d, ngroups, swz);
     else
          hipLaunchKernelGGL((k_dense_direct<K, R, GPT, false>), grid, dim3(kBlock), 0, s,
                             static_cast<V*>(psi), static_cast<const V*>(mat), d, ngroups, swz);
     return hipGetLastError();
}

template <typename R>
hipError_t dense_dispatch(void* psi, int n, const void* mat, const DenseDesc& d, hipStream_t s)
{
     switch (d.k) {
          case 1: return dense_k<1, R>(psi, n, mat, d, s);
          case 2: return dense_k<2, R>(psi, n, mat, d, s);
          case 3: return dense_k<3, R>(psi, n, mat, d, s);
          case 4: return dense_k<4, R>(psi, n, mat, d, s);
          case 5: return dense_k<5, R>(psi, n, mat, d, s);
          default: return hipErrorInvalidValue;
     }
}

InsertSet ctrl_insert_set(uint64_t ctrlmask)
{
     InsertSet s{};
     for (int p = 0; p < 64 && s.n < 63; ++p)
          if ((ctrlmask >> p) & 1) s.p[s.n++] = static_cast<uint8_t>(p);
     return s;
}

template <int K, typename R>
hipError_t diag_k(void* psi, int n, const double* dre, const DenseDesc& d, hipStream_t s)
{
     DiagVals<R> dv{};
     for (int i = 0; i < (2 << K); ++i) dv.v[i] = static_cast<R>(dre[i]);
     const InsertSet cins = ctrl_insert_set(d.ctrlmask);
     if (cins.n > n) return hipErrorInvalidValue;
     const uint64_t count = 1ull << (n - cins.n);
     const unsigned grid = grid_for(count, static_cast<uint64_t>(kBlock) * 4);
     // Non-temporal when the control bits leave whole 128-byte lines intact.
     const uint64_t line_mask = sizeof(VT<R>) == 16 ? 7 : 15;
     if ((d.ctrlmask & line_mask) == 0)
          hipLaunchKernelGGL((k_diag<K, R, true>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<VT<R>*>(psi), dv, cins, d.ctrlmask, d.pos[0], d.pos[1],
                             d.pos[2], d.pos[3], d.pos[4], count);
     else
          hipLaunchKernelGGL((k_diag<K, R, false>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<VT<R>*>(psi), dv, cins, d.ctrlmask, d.pos[0], d.pos[1],
                             d.pos[2], d.pos[3], d.pos[4], count);
     return hipGetLastError();
}

template <typename R>
hipError_t diag_dispatch(void* psi, int n, const double* dre, const DenseDesc& d, hipStream_t s)
{
     switch (d.k) {
          case 1: return diag_k<1, R>(psi, n, dre, d, s);
          case 2: return diag_k<2, R>(psi, n, dre, d, s);
          case 3: return diag_k<3, R>(psi, n, dre, d, s);
          case 4: return diag_k<4, R>(psi, n, dre, d, s);
          case 5: return diag_k<5, R>(psi, n, dre, d, s);
          default: return hipErrorInvalidValue;
     }
}

template <int NOUT, class Launch>
hipError_t reduce_launch(uint64_t work, double* scratch, double* out, hipStream_t s, Launch&& l)
{
     uint64_t g = (work + kBlock - 1) / kBlock;
     if (g < 1) g = 1;
     if (g > static_cast<uint64_t>(kReduceBlocks)) g = kReduceBlocks;
     const int grid = static_cast<int>(g);
     l(grid);
     hipError_t e = hipGetLastError();
     if (e != hipSuccess) return e;
     hipLaunchKernelGGL((k_reduce_final<NOUT>), dim3(1), dim3(kBlock), 0, s, scratch, grid, out);
     return hipGetLastError();
}

}  // namespace

namespace {
constexpr auto kRelaxed = std::memory_order_relaxed;
}  // namespace

void set_tile_high_target(int pos) { g_high_target.store(pos, kRelaxed); }
int get_tile_high_target() { return g_high_target.load(kRelaxed); }
void set_tile_high_target_wide(int pos) { g_high_target_wide.store(pos, kRelaxed); }
int get_tile_high_target_wide() { return g_high_target_wide.load(kRelaxed); }
int get_tile_high_target_for(int k) { return k == 4 ? get_tile_high_target_wide() : get_tile_high_target(); }
void set_block_swizzle(int bits) { g_swizzle_bits.store(bits < 0 ? 0 : (bits > 16 ? 16 : bits), kRelaxed); }
int get_block_swizzle() { return g_swizzle_bits.load(kRelaxed); }
void set_block_swizzle_wide(int bits) { g_swizzle_bits_wide.store(bits < 0 ? 0 : (bits > 16 ? 16 : bits), kRelaxed); }
int get_block_swizzle_wide() { return g_swizzle_bits_wide.load(kRelaxed); }
int get_block_swizzle_for(int k) { return k >= 4 ? get_block_swizzle_wide() : get_block_swizzle(); }
void set_tile_mfma(int mode) { g_tile_mfma.store(mode ? 1 : 0, kRelaxed); }
int get_tile_mfma() { return g_tile_mfma.load(kRelaxed); }
void set_dense5_tile(int on) { g_dense5_tile.store(on ? 1 : 0, kRelaxed); }
int get_dense5_tile() { return g_dense5_tile.load(kRelaxed); }
void set_direct_block(int bs) { g_direct_block.store(bs, kRelaxed); }
int get_direct_block() { return g_direct_block.load(kRelaxed); }
void set_permute_swizzle(int bits) { g_permute_swizzle.store(bits < 0 ? 0 : (bits > 16 ? 16 : bits), kRelaxed); }
int get_permute_swizzle() { return g_permute_swizzle.load(kRelaxed); }
void set_permute_block(int bs) { g_permute_block.store(bs == 512 ? 512 : 256, kRelaxed); }
int get_permute_block() { return g_permute_block.load(kRelaxed); }
void set_tile_threads(int nt) { g_tile_threads.store(nt == 256 || nt == 512 ? nt : 0, kRelaxed); }
int get_tile_threads() { return g_tile_threads.load(kRelaxed); }
int tile_threads_for(Prec p, int kmax)
{
     const int nt = get_tile_threads();
     return nt ? nt : (p == Prec::C64 || kmax >= 5 ? 256 : 512);
}

hipError_t launch_dense(void* psi, Prec prec, int n_qubits, const void* mat, const DenseDesc& d,
                        hipStream_t s)
{
     return prec == Prec::C128 ? dense_dispatch<double>(psi, n_qubits, mat, d, s)
                               : dense_dispatch<float>(psi, n_qubits, mat, d, s);
}

hipError_t launch_diag(void* psi, Prec prec, int n_qubits, const double* diag_re_im,
                       const DenseDesc& d, hipStream_t s)
{
     return prec == Prec::C128 ? diag_dispatch<double>(psi, n_qubits, diag_re_im, d, s)
                               : diag_dispatch<float>(psi, n_qubits, diag_re_im, d, s);
}

hipError_t launch_scale(void* psi, Prec prec, uint64_t n, double fre, double fim, hipStream_t s)
{
     const unsigned grid = grid_for(n, kBlock * 4);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_scale<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(psi), fre, fim, n);
     else
          hipLaunchKernelGGL((k_scale<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(psi), static_cast<float>(fre),
                             static_cast<float>(fim), n);
     return hipGetLastError();
}

hipError_t launch_fill(void* psi, Prec prec, uint64_t n, double vre, double vim, hipStream_t s)
{
     const unsigned grid = grid_for(n, kBlock * 4);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_fill<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(psi), vre, vim, n);
     else
          hipLaunchKernelGGL((k_fill<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(psi), static_cast<float>(vre),
                             static_cast<float>(vim), n);
     return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, Prec prec, uint64_t n, hipStream_t s)
{
     const uint64_t bytes = n * amp_bytes(prec);
     const uint64_t words = bytes / 16;
     if (words > 0) {
          const unsigned grid = grid_for(words, kBlock * 4);
          hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(kBlock), 0, s, static_cast<uint4*>(dst),
                             static_cast<const uint4*>(src), words);
          hipError_t e = hipGetLastError();
          if (e != hipSuccess) return e;
     }
     const uint64_t tail = bytes - words * 16;
     if (tail)
          return hipMemcpyAsync(static_cast<char*>(dst) + words * 16,
                                static_cast<const char*>(src) + words * 16, tail,
                                hipMemcpyDeviceToDevice, s);
     return hipSuccess;
}

hipError_t launch_normalize(void* psi, Prec prec, uint64_t n, uint64_t msk, uint64_t val, double f,
                            hipStream_t s)
{
     const unsigned grid = grid_for(n, kBlock * 4);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_normalize<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(psi), n, msk, val, f);
     else
          hipLaunchKernelGGL((k_normalize<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(psi), n, msk, val, static_cast<float>(f));
     return hipGetLastError();
}

hipError_t launch_remove_bit(void* dst, const void* src, Prec prec, uint64_t n_src, int pos,
                             int keep, hipStream_t s)
{
     const uint64_t n_dst = n_src / 2;
     const unsigned grid = grid_for(n_dst, kBlock * 4);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_remove_bit<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(dst), static_cast<const double2*>(src), n_dst,
                             pos, static_cast<uint64_t>(keep));
     else
          hipLaunchKernelGGL((k_remove_bit<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(dst), static_cast<const float2*>(src), n_dst, pos,
                             static_cast<uint64_t>(keep));
     return hipGetLastError();
}

hipError_t launch_permute(void* dst, const void* src, Prec prec, int n_qubits,
                          const uint8_t* src_pos, hipStream_t s)
{
     if (n_qubits > 40) return hipErrorInvalidValue;
     const uint64_t n = 1ull << n_qubits;
     if (n_qubits >= 12) {
          // Tile bits: dst bits 0..5, the dst positions of src bits 0..5, then
          // the lowest remaining dst bits up to 12.
          int d_of[40];
          for (int p = 0; p < n_qubits; ++p) d_of[src_pos[p]] = p;
          uint64_t U = 63;
          for (int b = 0; b < 6; ++b) U |= 1ull << d_of[b];
          for (int p = 0; p < n_qubits && __builtin_popcountll(U) < 12; ++p) U |= 1ull << p;
          const int T = __builtin_popcountll(U);
          PermTile P{};
          P.nq = n_qubits;
          int rank_u[40];
          for (int p = 0; p < n_qubits; ++p) {
               P.src_pos[p] = src_pos[p];
               if ((U >> p) & 1) {
                    rank_u[p] = P.u.n;
                    P.u.p[P.u.n++] = static_cast<uint8_t>(p);
               }
          }
          uint64_t S = 0;
          for (int p = 0; p < n_qubits; ++p)
               if ((U >> p) & 1) S |= 1ull << src_pos[p];
          int a = 0;
          for (int b = 0; b < n_qubits; ++b)
               if ((S >> b) & 1) {
                    P.src_sorted[a] = static_cast<uint8_t>(b);
                    P.qbit_of_rbit[a] = static_cast<uint8_t>(rank_u[d_of[b]]);
                    ++a;
               }
          const uint64_t nblocks = n >> T;
          const dim3 grid = grid_2d(nblocks);
          int swz = g_permute_swizzle.load(std::memory_order_relaxed);
          while (swz > 0 && (nblocks % (1ull << swz)) != 0) --swz;
          const bool wide = g_permute_block.load(std::memory_order_relaxed) == 512;
#define HQ_PERM_TILE(RT, VT_, TT)                                                                  \
     do {                                                                                         \
          if (wide)                                                                               \
               hipLaunchKernelGGL((k_permute_tile<RT, TT, 9>), grid, dim3(512), 0, s,             \
                                  static_cast<VT_*>(dst), static_cast<const VT_*>(src), P, swz);   \
          else                                                                                    \
               hipLaunchKernelGGL((k_permute_tile<RT, TT, 8>), grid, dim3(256), 0, s,             \
                                  static_cast<VT_*>(dst), static_cast<const VT_*>(src), P, swz);   \
     } while (0)
          if (prec == Prec::C128) {
               if (T == 10) HQ_PERM_TILE(double, double2, 10);
               else if (T == 11) HQ_PERM_TILE(double, double2, 11);
               else HQ_PERM_TILE(double, double2, 12);
          }
          else {
               if (T == 10) HQ_PERM_TILE(float, float2, 10);
               else if (T == 11) HQ_PERM_TILE(float, float2, 11);
               else HQ_PERM_TILE(float, float2, 12);
          }
#undef HQ_PERM_TILE
          return hipGetLastError();
     }
     InsertSet sp{};
     sp.n = n_qubits;
     for (int p = 0; p < n_qubits; ++p) sp.p[p] = src_pos[p];
     const unsigned grid = grid_for(n, kBlock * 8);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_permute<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(dst), static_cast<const double2*>(src), n,
                             n_qubits, sp);
     else
          hipLaunchKernelGGL((k_permute<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(dst), static_cast<const float2*>(src), n,
                             n_qubits, sp);
     return hipGetLastError();
}

hipError_t launch_pauli_accumulate(void* dst, const void* src, Prec prec, uint64_t n,
                                   uint64_t xmask, uint64_t zmask, double cre, double cim,
                                   hipStream_t s)
{
     const unsigned grid = grid_for(n, kBlock * 4);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_pauli_acc<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(dst), static_cast<const double2*>(src), n,
                             xmask, zmask, cre, cim);
     else
          hipLaunchKernelGGL((k_pauli_acc<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(dst), static_cast<const float2*>(src), n, xmask,
                             zmask, static_cast<float>(cre), static_cast<float>(cim));
     return hipGetLastError();
}

hipError_t launch_masked_norm(const void* psi, Prec prec, int n_qubits, uint64_t msk, uint64_t val,
                              double* scratch, double* out, hipStream_t s)
{
     const InsertSet ins = ctrl_insert_set(msk);
     if (ins.n > n_qubits) return hipErrorInvalidValue;
     const uint64_t count = 1ull << (n_qubits - ins.n);
     return reduce_launch<1>(count, scratch, out, s, [&](int grid) {
          if (prec == Prec::C128)
               hipLaunchKernelGGL((k_masked_norm<double>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const double2*>(psi), count, ins, val, scratch);
          else
               hipLaunchKernelGGL((k_masked_norm<float>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const float2*>(psi), count, ins, val, scratch);
     });
}

hipError_t launch_two_bucket(const void* psi, Prec prec, int n_qubits, int pos, double* scratch,
                             double* out, hipStream_t s)
{
     const uint64_t n = 1ull << n_qubits;
     return reduce_launch<2>(n, scratch, out, s, [&](int grid) {
          if (prec == Prec::C128)
               hipLaunchKernelGGL((k_two_bucket<double>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const double2*>(psi), n, pos, scratch);
          else
               hipLaunchKernelGGL((k_two_bucket<float>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const float2*>(psi), n, pos, scratch);
     });
}

hipError_t launch_entropy(const void* psi, Prec prec, uint64_t n, double* scratch, double* out,
                          hipStream_t s)
{
     return reduce_launch<1>(n, scratch, out, s, [&](int grid) {
          if (prec == Prec::C128)
               hipLaunchKernelGGL((k_entropy<double>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const double2*>(psi), n, scratch);
          else
               hipLaunchKernelGGL((k_entropy<float>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const float2*>(psi), n, scratch);
     });
}

hipError_t launch_pauli_expectation(const void* psi, Prec prec, uint64_t n, uint64_t xmask,
                                    uint64_t zmask, double* scratch, double* out, hipStream_t s)
{
     return reduce_launch<2>(n, scratch, out, s, [&](int grid) {
          if (prec == Prec::C128)
               hipLaunchKernelGGL((k_pauli_exp<double>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const double2*>(psi), n, xmask, zmask, scratch);
          else
               hipLaunchKernelGGL((k_pauli_exp<float>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const float2*>(psi), n, xmask, zmask, scratch);
     });
}

hipError_t launch_random_state(void* psi, Prec prec, uint64_t n, uint64_t seed, hipStream_t s)
{
     const unsigned grid = grid_for(n, kBlock * 4);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_random_gaussian<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<double2*>(psi), n, seed);
     else
          hipLaunchKernelGGL((k_random_gaussian<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<float2*>(psi), n, seed);
     return hipGetLastError();
}

hipError_t launch_compare(const void* psi, const void* ref, Prec prec, uint64_t n, double* scratch,
                          double* out, hipStream_t s)
{
     uint64_t g = (n + kBlock - 1) / kBlock;
     if (g < 1) g = 1;
     if (g > static_cast<uint64_t>(kReduceBlocks)) g = kReduceBlocks;
     const int grid = static_cast<int>(g);
     if (prec == Prec::C128)
          hipLaunchKernelGGL((k_compare<double>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<const double2*>(psi), static_cast<const double2*>(ref), n,
                             scratch);
     else
          hipLaunchKernelGGL((k_compare<float>), dim3(grid), dim3(kBlock), 0, s,
                             static_cast<const float2*>(psi), static_cast<const float2*>(ref), n,
                             scratch);
     hipError_t e = hipGetLastError();
     if (e != hipSuccess) return e;
     hipLaunchKernelGGL(k_compare_final, dim3(1), dim3(kBlock), 0, s, scratch, grid, out);
     return hipGetLastError();
}

hipError_t launch_block_sums(const void* psi, Prec prec, uint64_t n, uint64_t nb, double* out,
                             hipStream_t s)
{
     if (nb == 0 || n % nb != 0) return hipErrorInvalidValue;
     const uint64_t K = n / nb;
     if (K >= kBlock) {
          if (prec == Prec::C128)
               hipLaunchKernelGGL((k_block_sums_wg<double>), dim3(static_cast<unsigned>(nb)),
                                  dim3(kBlock), 0, s, static_cast<const double2*>(psi), K, out);
          else
               hipLaunchKernelGGL((k_block_sums_wg<float>), dim3(static_cast<unsigned>(nb)),
                                  dim3(kBlock), 0, s, static_cast<const float2*>(psi), K, out);
     }
     else {
          const unsigned grid = static_cast<unsigned>((nb + kBlock - 1) / kBlock);
          if (prec == Prec::C128)
               hipLaunchKernelGGL((k_block_sums_seq<double>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const double2*>(psi), K, nb, out);
          else
               hipLaunchKernelGGL((k_block_sums_seq<float>), dim3(grid), dim3(kBlock), 0, s,
                                  static_cast<const float2*>(psi), K, nb, out);
     }
     return hipGetLastError();
}

}  // namespace hq
