// Device helpers shared by the gfx950 kernel translation units: complex
// element types, zero-bit insertion, non-temporal accesses, the per-group
// fused matrix-vector product and launch-shape helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace hq {
namespace {

constexpr int kBlock = 256;
constexpr uint64_t kMaxGrid = 1ull << 20;

template <typename R>
struct C2;
template <>
struct C2<double> {
     using T = double2;
};
template <>
struct C2<float> {
     using T = float2;
};

template <typename R>
using VT = typename C2<R>::T;

__device__ __forceinline__ uint64_t insert_zeros(uint64_t t, const InsertSet& s)
{
     for (int i = 0; i < s.n; ++i) {
          const uint64_t lo = t & ((1ull << s.p[i]) - 1ull);
          t = ((t ^ lo) << 1) | lo;
     }
     return t;
}

inline unsigned grid_for(uint64_t work_items, uint64_t per_block)
{
     uint64_t g = (work_items + per_block - 1) / per_block;
     if (g == 0) g = 1;
     if (g > kMaxGrid) g = kMaxGrid;
     return static_cast<unsigned>(g);
}

// Non-temporal 16-byte / 8-byte global accesses for streamed amplitudes.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2 ld_nt(const double2* p)
{
     const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
     return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ld_nt(const float2* p)
{
     const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
     return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(double2* p, double2 a)
{
     const d2v v = {a.x, a.y};
     __builtin_nontemporal_store(v, reinterpret_cast<d2v*>(p));
}
__device__ __forceinline__ void st_nt(float2* p, float2 a)
{
     const f2v v = {a.x, a.y};
     __builtin_nontemporal_store(v, reinterpret_cast<f2v*>(p));
}
template <bool NT, typename V>
__device__ __forceinline__ V ld(const V* p)
{
     return NT ? ld_nt(p) : *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void st(V* p, V a)
{
     if (NT) st_nt(p, a);
     else *p = a;
}

// out_i = sum_j m[i][j] v_j over one group held in registers; `put(i, o)`
// stores output i.  The matrix is read row by row with uniform loads.
template <int K, typename V, typename R, typename Put>
__device__ __forceinline__ void group_matvec(const V* __restrict__ mat, const V (&v)[1 << K], Put&& put)
{
     constexpr int D = 1 << K;
#pragma unroll
     for (int i = 0; i < D; ++i) {
          R re = 0, im = 0;
#pragma unroll
          for (int j = 0; j < D; ++j) {
               const V m = mat[i * D + j];
               re = fma(m.x, v[j].x, re);
               re = fma(-m.y, v[j].y, re);
               im = fma(m.x, v[j].y, im);
               im = fma(m.y, v[j].x, im);
          }
          V o;
          o.x = re;
          o.y = im;
          put(i, o);
     }
}

__device__ __forceinline__ uint64_t block_linear()
{
     return static_cast<uint64_t>(blockIdx.y) * gridDim.x + blockIdx.x;
}

// Grid covering `blocks` blocks exactly (a power of two or not) with a 2-D
// launch so that very large states stay within the 32-bit dispatch limits.
inline dim3 grid_2d(uint64_t blocks)
{
     if (blocks == 0) blocks = 1;
     const uint64_t gx = blocks < 65536 ? blocks : 65536;
     return dim3(static_cast<unsigned>(gx), static_cast<unsigned>((blocks + gx - 1) / gx));
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): counter-based, so amplitude i draws from
// counter i independently of the launch shape.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
     constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
     constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
     for (int r = 0; r < 10; ++r) {
          const uint64_t p0 = static_cast<uint64_t>(M0) * c.x;
          const uint64_t p1 = static_cast<uint64_t>(M1) * c.z;
          const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
          const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
          c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
          k.x += W0;
          k.y += W1;
     }
     return c;
}

// LDS slot of tile element r: XOR-swizzled so that both the chunk-order
// accesses (consecutive r across lanes) and the group-order accesses (lanes
// strided by powers of two) spread over the banks.
template <typename V>
__device__ __forceinline__ int tile_slot(int r)
{
     constexpr int S = sizeof(V) == 16 ? 4 : 5;
     return r ^ ((r >> S) & ((1 << S) - 1));
}

// Block order of the streaming passes.  Blocks start in index order, so with
// the identity mapping the blocks in flight at one time cover one narrow
// address range inside each of the 2^k chunk strides of a pass; measured on
// MI355X this concentrates the traffic on few HBM channels when the strides
// are large (targets at bits 18-29: 4.0-4.6 TB/s for a 30-qubit k = 4 pass
// against 6.0-6.6 TB/s at low bits).  Reversing the low `bits` bits of the
// block index spreads consecutive blocks over 2^bits far-apart address
// ranges; it is a bijection of [0, nblocks) when nblocks is a multiple of
// 2^bits.
__device__ __forceinline__ uint64_t swizzle_block(uint64_t b, int bits)
{
     if (bits <= 0) return b;
     const uint64_t mask = (1ull << bits) - 1ull;
     const uint32_t r = __builtin_bitreverse32(static_cast<uint32_t>(b & mask)) >> (32 - bits);
     return (b & ~mask) | r;
}

// 53-bit uniform in [0, 1) from two 32-bit Philox outputs.
__device__ __forceinline__ double uniform53(uint32_t hi, uint32_t lo)
{
     return static_cast<double>((static_cast<uint64_t>(hi) << 21) ^ lo) * 0x1p-53;
}

}  // namespace
}  // namespace hq
