// cull_math.hpp -- device-side FP32 math of the culling path, in the exact
// evaluation order of nvpro-pipeline's dp::math and dp::culling::cpu code.
//
// Every function is a sequence of separate IEEE binary32 multiplies and adds.
// The translation units that include this header are compiled with
// -ffp-contract=off and the pragma below, so no multiply-add is fused and
// nothing is reassociated; float denormals are preserved
// (-fno-gpu-flush-denormals-to-zero).  With that, the device produces the
// same bits as the reference's scalar C++ path.
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace dpc {

// A 4x4 row-major matrix passed by value in kernel arguments: the compiler
// loads it into scalar registers once per wave, and the VALU multiplies read
// it from there.
struct Mat16 {
  float m[16];
};

__device__ __forceinline__ float4 add4(float4 a, float4 b)
{
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__device__ __forceinline__ float4 scale4(float s, float4 v)
{
  return make_float4(v.x * s, v.y * s, v.z * s, v.w * s);
}

// v * M, dp/math/Matmnt.h:1371-1379: out[c] = ((v0*M0c + v1*M1c) + v2*M2c) + v3*M3c.
__device__ __forceinline__ float4 vecmat(float4 v, const float* m)
{
  float4 r;
  r.x = v.x * m[0];
  r.x = r.x + v.y * m[4];
  r.x = r.x + v.z * m[8];
  r.x = r.x + v.w * m[12];
  r.y = v.x * m[1];
  r.y = r.y + v.y * m[5];
  r.y = r.y + v.z * m[9];
  r.y = r.y + v.w * m[13];
  r.z = v.x * m[2];
  r.z = r.z + v.y * m[6];
  r.z = r.z + v.z * m[10];
  r.z = r.z + v.w * m[14];
  r.w = v.x * m[3];
  r.w = r.w + v.y * m[7];
  r.w = r.w + v.z * m[11];
  r.w = r.w + v.w * m[15];
  return r;
}

// Row i of A*B given row i of A and the four rows of B
// (dp/math/Matmnt.h:1381-1417: ((a_i0*B0 + a_i1*B1) + a_i2*B2) + a_i3*B3, per column).
__device__ __forceinline__ float4 rowmat(float4 a, float4 b0, float4 b1, float4 b2, float4 b3)
{
  float4 r;
  r.x = a.x * b0.x;
  r.x = r.x + a.y * b1.x;
  r.x = r.x + a.z * b2.x;
  r.x = r.x + a.w * b3.x;
  r.y = a.x * b0.y;
  r.y = r.y + a.y * b1.y;
  r.y = r.y + a.z * b2.y;
  r.y = r.y + a.w * b3.y;
  r.z = a.x * b0.z;
  r.z = r.z + a.y * b1.z;
  r.z = r.z + a.z * b2.z;
  r.z = r.z + a.w * b3.z;
  r.w = a.x * b0.w;
  r.w = r.w + a.y * b1.w;
  r.w = r.w + a.z * b2.w;
  r.w = r.w + a.w * b3.w;
  return r;
}

// Object-space box transformed to world space:
// GroupCPU::updateOBBs (dp/culling/cpu/src/ManagerImpl.cpp:131, 152-156).
struct Obb {
  float4 p, ex, ey, ez;
};

// lower = (lo.xyz, 1) and extent = (ex.xyz, 0), as ManagerBitSet::objectSetBoundingBox
// stores them (dp/culling/src/ManagerBitSet.cpp:99-109); r0..r3 are the rows of the
// object's matrix.
__device__ __forceinline__ Obb make_obb(float4 box_a, float2 box_b, float4 r0, float4 r1, float4 r2, float4 r3)
{
  Obb o;
  const float lx = box_a.x, ly = box_a.y, lz = box_a.z;
  // point = (lx, ly, lz, 1) * M, left to right.
  o.p.x = lx * r0.x;
  o.p.x = o.p.x + ly * r1.x;
  o.p.x = o.p.x + lz * r2.x;
  o.p.x = o.p.x + 1.0f * r3.x;
  o.p.y = lx * r0.y;
  o.p.y = o.p.y + ly * r1.y;
  o.p.y = o.p.y + lz * r2.y;
  o.p.y = o.p.y + 1.0f * r3.y;
  o.p.z = lx * r0.z;
  o.p.z = o.p.z + ly * r1.z;
  o.p.z = o.p.z + lz * r2.z;
  o.p.z = o.p.z + 1.0f * r3.z;
  o.p.w = lx * r0.w;
  o.p.w = o.p.w + ly * r1.w;
  o.p.w = o.p.w + lz * r2.w;
  o.p.w = o.p.w + 1.0f * r3.w;
  o.ex = scale4(box_a.w, r0);
  o.ey = scale4(box_b.x, r1);
  o.ez = scale4(box_b.y, r2);
  return o;
}

// determineCullFlags (ManagerImpl.cpp:199-229) folded into the AND over the
// eight corners.  Per axis the reference sets the "low" flag when c <= -w and,
// only otherwise, the "high" flag when w <= c.  Over the corners:
//   low bit common to all corners   = AND_k (c_k <= -w_k)
//   high bit common to all corners  = AND_k (!(c_k <= -w_k) && w_k <= c_k)
//                                   = !(OR_k (c_k <= -w_k)) && AND_k (w_k <= c_k)
// so two compares per axis and corner suffice (48 per test), with the AND/OR
// accumulation carried in wave-wide lane masks.
struct PlaneAcc {
  bool lo_all[3], lo_any[3], hi_all[3];
};

__device__ __forceinline__ void corner_flags(float4 c, PlaneAcc& s)
{
  const float cv[3] = {c.x, c.y, c.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const bool lo = cv[a] <= -c.w;
    const bool hi = c.w <= cv[a];
    // bitwise (not short-circuit) operators: straight-line code, no divergent branches
    s.lo_all[a] = s.lo_all[a] & lo;
    s.lo_any[a] = s.lo_any[a] | lo;
    s.hi_all[a] = s.hi_all[a] & hi;
  }
}

// isVisible(projection, OBB) (ManagerImpl.cpp:263-289).  The reference returns
// !cfo || !cfa; since cfa is the AND and cfo the OR of the same eight flag words,
// cfo == 0 implies cfa == 0, so the result is cfa == 0.
__device__ __forceinline__ bool obb_visible(const Obb& o, const float* vp)
{
  const float4 v0 = vecmat(o.p, vp);
  const float4 x = vecmat(o.ex, vp);
  const float4 y = vecmat(o.ey, vp);
  const float4 z = vecmat(o.ez, vp);
  const float4 c1 = add4(v0, x);
  const float4 c2 = add4(v0, y);
  const float4 c3 = add4(c1, y);
  const float4 c4 = add4(v0, z);
  const float4 c5 = add4(c1, z);
  const float4 c6 = add4(c2, z);
  const float4 c7 = add4(c3, z);
  PlaneAcc s{{true, true, true}, {false, false, false}, {true, true, true}};
  corner_flags(v0, s);
  corner_flags(c1, s);
  corner_flags(c2, s);
  corner_flags(c3, s);
  corner_flags(c4, s);
  corner_flags(c5, s);
  corner_flags(c6, s);
  corner_flags(c7, s);
  bool culled = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) culled = culled | s.lo_all[a] | (s.hi_all[a] & !s.lo_any[a]);
  return !culled;
}

// The raw inputs of one object as loaded from the SoA arrays: box (lower.xyz,
// extent.x), (extent.y, extent.z) and the four matrix rows.  Loading and using
// are separate steps so that kernels can issue the loads of the next tile
// before they compute the current one.
struct RawObj {
  float4 a;
  float2 b;
  float4 r0, r1, r2, r3;
};

// Loads through a pointer known to address global memory.  Pointers read from
// a table in memory (a scene descriptor) are generic to the compiler, and a
// generic (flat) load makes it wait for every load in flight before using any
// of them, which would also wait for the prefetch of the next tile.
__device__ __forceinline__ float4 gload(const float4* p)
{
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = *(const __attribute__((address_space(1))) v4*)(reinterpret_cast<uintptr_t>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float2 gload(const float2* p)
{
  typedef float v2 __attribute__((ext_vector_type(2)));
  const v2 v = *(const __attribute__((address_space(1))) v2*)(reinterpret_cast<uintptr_t>(p));
  return make_float2(v.x, v.y);
}

__device__ __forceinline__ uint32_t gload(const uint32_t* p)
{
  return *(const __attribute__((address_space(1))) uint32_t*)(reinterpret_cast<uintptr_t>(p));
}

// Loads object i with matrix m.  The loads are unconditional so that no branch
// separates them from the compute (which lets the compiler wait for exactly
// the loads a tile needs): the caller passes indices that are in range, and
// keeps its own per-lane flags for lanes without an object or with a matrix
// index outside the array.
__device__ __forceinline__ RawObj load_raw(const float4* __restrict__ box_a, const float2* __restrict__ box_b,
                                           const float4* __restrict__ planes, size_t ld, size_t i, size_t m)
{
  RawObj r;
  r.a = gload(box_a + i);
  r.b = gload(box_b + i);
  r.r0 = gload(planes + m);
  r.r1 = gload(planes + ld + m);
  r.r2 = gload(planes + 2 * ld + m);
  r.r3 = gload(planes + 3 * ld + m);
  return r;
}

// Streaming loads (read once, not kept in the caches) for data that one launch
// reads exactly once.
__device__ __forceinline__ float4 load_stream(const float4* p)
{
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float2 load_stream(const float2* p)
{
  typedef float v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return make_float2(v.x, v.y);
}

// load_raw with streaming loads for the boxes and, when each object has a
// matrix of its own (StreamMatrices), for the matrix rows too; a matrix shared
// by several objects (an indexed scene) is read through the caches.
template <bool StreamMatrices>
__device__ __forceinline__ RawObj load_raw_stream(const float4* __restrict__ box_a, const float2* __restrict__ box_b,
                                                  const float4* __restrict__ planes, size_t ld, size_t i, size_t m)
{
  RawObj r;
  r.a = load_stream(box_a + i);
  r.b = load_stream(box_b + i);
  if (StreamMatrices) {
    r.r0 = load_stream(planes + m);
    r.r1 = load_stream(planes + ld + m);
    r.r2 = load_stream(planes + 2 * ld + m);
    r.r3 = load_stream(planes + 3 * ld + m);
  } else {
    r.r0 = planes[m];
    r.r1 = planes[ld + m];
    r.r2 = planes[2 * ld + m];
    r.r3 = planes[3 * ld + m];
  }
  return r;
}

// Per-lane object slot of a tile: the (clamped) indices to load from and what
// the lane's visibility bit becomes.  An object whose matrix index lies outside
// the matrix array reads nothing of it and is reported visible (the value the
// reference's test gives an object whose clip coordinates are NaN: no flag set).
struct Slot {
  size_t i, m;
  bool valid;       // the lane holds an object
  bool bad_matrix;  // its matrix index is outside the array
};

__device__ __forceinline__ Slot make_slot(size_t i, size_t n, const uint32_t* __restrict__ xform, size_t ld)
{
  Slot s;
  s.valid = i < n;
  s.i = s.valid ? i : 0;
  const size_t m = xform ? static_cast<size_t>(gload(xform + s.i)) : s.i;
  s.bad_matrix = m >= ld;
  s.m = s.bad_matrix ? 0 : m;
  return s;
}

__device__ __forceinline__ bool slot_visible(const Slot& s, bool test) { return s.valid & (s.bad_matrix | test); }

__device__ __forceinline__ Obb obb_of(const RawObj& r) { return make_obb(r.a, r.b, r.r0, r.r1, r.r2, r.r3); }

}  // namespace dpc
