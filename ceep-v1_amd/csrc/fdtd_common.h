// Shared device-side definitions of the batched FDTD engine (gfx950 / CDNA4).
//
// Data layout in HBM (all FP32):
//   field[b][a][c]      2D: b = TX (batch), a = x - n, c = y - n        (c fastest)
//   field[b][a][c][d]   3D: ... d = z - n                               (d fastest)
// over the "cavity box" [n, N-n] on every axis.  With every source strictly
// inside the E-update interior, the reference's CPML auxiliaries and every
// field outside that box stay exactly zero (the profile grading makes
// c[n-1] = b[n-1] - 1 = 0), so only the box is stored and stepped.
// Rows are padded to `pitch` floats (a multiple of 64 -> 256 B aligned rows).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ceep {

constexpr int kMaxPoles = 2;          // Debye + Lorentz per cell
constexpr int kMaxBatch = 1024;
constexpr int kMaxCpml = 64;          // physical-CPML mode: largest slab thickness

// One ADE pole's per-cell coefficient maps (FP32, cavity box layout without
// the batch axis).  k2 is pre-divided by dt: J += k2dt * (E_new - E_old).
struct PoleMaps {
  const float* k1;
  const float* k2dt;
  const float* k3;     // Lorentz only (J_prev coefficient); null for Debye
};

// Raw buffer addressing (CDNA buffer instructions).  A plane is described by
// a 128-bit resource (base, size in bytes; word 3 = 0x00020000, the gfx9
// family's raw 32-bit format); a lane addresses it with a 32-bit byte offset.
// The hardware range check turns any offset >= the size into a load of 0
// and a dropped store, so out-of-window cells need neither a branch nor a
// select: their offset is pushed to kOffNone or beyond (planes stay below
// 2 GiB, so kOffNone plus a row offset does not wrap).
constexpr unsigned kOffNone = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* base,
                                                             unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, 0);
}

// Raw buffer access with a uniform SGPR offset (slab, set or component) on
// top of the lane's VGPR offset; constant component offsets added to the
// VGPR offset fold into the instruction.
__device__ __forceinline__ float rload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void rstore(float v, __amdgpu_buffer_rsrc_t r, unsigned voff,
                                       unsigned soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ int rload_short(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return (int)(short)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}

// DPP wave shifts: lane l receives lane l+1 (0 for lane 63), resp. lane l-1
// (0 for lane 0)
__device__ __forceinline__ float lane_next(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_prev(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

// Workgroup barrier for LDS exchanges: waits for this wavefront's LDS
// operations only, not for its global loads in flight (prefetches)
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The 2D TMz cell updates with their roundings fixed: every 2D forward kernel
// (one step per launch, one- and multi-wavefront windows, uneven rounds)
// calls these, so their results are bit-identical whatever the compiler
// would contract on its own in each instantiation.
//   H: h - ch d (hx), h + ch d (hy), each one fused multiply-add
__device__ __forceinline__ float h_minus(float h, float ch, float d) {
  return __builtin_fmaf(-ch, d, h);
}
__device__ __forceinline__ float h_plus(float h, float ch, float d) {
  return __builtin_fmaf(ch, d, h);
}
//   E: ca e0 + cb (curl idx)
__device__ __forceinline__ float e_step(float ca, float cb, float e0, float curl, float idx) {
  return __builtin_fmaf(ca, e0, cb * (curl * idx));
}
//   E with the Debye current: ca e0 + cb (curl idx) - cb sgj, sgj = (1 + k1)/2 J
__device__ __forceinline__ float e_step_debye(float ca, float cb, float e0, float curl, float idx,
                                              float k1, float jd) {
  const float sgj = 0.5f * (1.f + k1) * jd;
  return __builtin_fmaf(-cb, sgj, __builtin_fmaf(ca, e0, cb * (curl * idx)));
}
//   Debye current: k1 J + k2dt (e_new - e_old)
__device__ __forceinline__ float j_debye(float k1, float k2dt, float jd, float de) {
  return __builtin_fmaf(k1, jd, k2dt * de);
}

template <int V>
struct Phase {
  static constexpr int value = V;
};

// Workgroups of a 1D grid are dealt round-robin over MI355X's 8 XCDs (each
// with its own L2).  With `on`, physical block b is renumbered so that XCD x
// runs one contiguous range of logical blocks: neighbouring tiles, which
// re-read each other's halo rows, then share an L2.  A bijection for any
// block count.
__device__ __forceinline__ int xcd_block(int on) {
  const int phys = blockIdx.x;
  if (!on) return phys;
  const int n = gridDim.x;
  const int q = n / 8, r = n % 8, xcd = phys % 8;
  return xcd * q + min(xcd, r) + phys / 8;
}

}  // namespace ceep
