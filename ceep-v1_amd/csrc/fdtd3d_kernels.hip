// Batched 3D Yee step kernels for gfx950.
//
// H kernel: the three H components in place (fdtd_3d_batched.py:753-818, on
// the cavity box, so the reference's H-CPML is the identity there).
// E kernel: the three E components in place on the E interior
// (fdtd_3d_batched.py:820-884) with the ADE of fdtd_2d.py:163-178 per
// component, then Ez injection (:949-951) and Ez probes (:956-958).
//
// Mapping: a 64x4 workgroup covers 64 consecutive z-cells (the contiguous
// axis) and 4 y-rows of one x-slab; blockIdx.z enumerates x-slabs times batch
// chunks, and each thread loops over BB batch planes so material maps are
// read once per chunk.
#include "fdtd_common.h"
#include "fdtd3d_kernels.h"

namespace ceep {

template <bool MAGNETIC, int BB>
__global__ __launch_bounds__(256) void h3d_kernel(Step3DArgs p) {
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z % p.Mx;
  const int chunk = blockIdx.z / p.Mx;
  if (b >= p.My || d >= p.Mz) return;
  const long long cell = (long long)a * p.sx + (long long)b * p.pitch + d;
  const bool xa = a < p.Mx - 1, yb = b < p.My - 1, zd = d < p.Mz - 1;
  float da = 1.f, c = p.ch;
  if (MAGNETIC) { da = p.da[cell]; c = p.dbi[cell]; }
  const int b0 = chunk * BB;
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    const int bt = b0 + k;
    if (bt >= p.B) break;
    const long long o = (long long)bt * p.bstride + cell;
    const float ex0 = p.ex[o], ey0 = p.ey[o], ez0 = p.ez[o];
    // hx: y < My-1, z < Mz-1 ; hy: x < Mx-1, z < Mz-1 ; hz: x < Mx-1, y < My-1
    if (yb && zd) {
      const float t = (p.ey[o + 1] - ey0) - (p.ez[o + p.pitch] - ez0);
      p.hx[o] = MAGNETIC ? da * p.hx[o] + c * t : p.hx[o] + c * t;
    }
    if (xa && zd) {
      const float t = (p.ez[o + p.sx] - ez0) - (p.ex[o + 1] - ex0);
      p.hy[o] = MAGNETIC ? da * p.hy[o] + c * t : p.hy[o] + c * t;
    }
    if (xa && yb) {
      const float t = (p.ex[o + p.pitch] - ex0) - (p.ey[o + p.sx] - ey0);
      p.hz[o] = MAGNETIC ? da * p.hz[o] + c * t : p.hz[o] + c * t;
    }
  }
}

// E update of one component at one cell: curl is the undivided difference
// (times 1/dx here), cb the FP32 (dt/eps)/(1+s) map value.
template <int POLES>
__device__ __forceinline__ float ade_update(const Step3DArgs& p, int comp, long long o,
                                            float e0, float ca, float cb, float curl,
                                            float dk1, float dk2, float lk1, float lk2,
                                            float lk3) {
  if (POLES == 0) return ca * e0 + cb * (curl * p.idx);
  float jd = 0.f, jl = 0.f, jlp = 0.f, sgj = 0.f;
  if (POLES & 1) { jd = p.jd[comp][o]; sgj += 0.5f * (1.f + dk1) * jd; }
  if (POLES & 2) { jl = p.jl_cur[comp][o]; jlp = p.jl_prev[comp][o]; sgj += jl; }
  const float en = ca * e0 + cb * (curl * p.idx) - cb * sgj;
  const float de = en - e0;
  if (POLES & 1) p.jd[comp][o] = dk1 * jd + dk2 * de;
  if (POLES & 2) p.jl_prev[comp][o] = lk1 * jl + lk3 * jlp + lk2 * de;
  return en;
}

template <int POLES, int BB>
__global__ __launch_bounds__(256) void e3d_kernel(Step3DArgs p) {
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z % p.Mx;
  const int chunk = blockIdx.z / p.Mx;
  if (b >= p.My || d >= p.Mz) return;
  const long long cell = (long long)a * p.sx + (long long)b * p.pitch + d;
  const bool interior = a >= 1 && a <= p.Mx - 2 && b >= 1 && b <= p.My - 2 &&
                        d >= 1 && d <= p.Mz - 2;
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
  if (!interior && probe0 < 0) return;
  float ca = 0.f, cb = 0.f, dk1 = 0.f, dk2 = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
  if (interior) {
    ca = p.ca[cell];
    cb = p.cb[cell];
    if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; }
    if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
  }
  const int b0 = chunk * BB;
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    const int bt = b0 + k;
    if (bt >= p.B) break;
    const long long o = (long long)bt * p.bstride + cell;
    float ezn = p.ez[o];
    if (interior) {
      const float hx0 = p.hx[o], hy0 = p.hy[o], hz0 = p.hz[o];
      const float cx = (hz0 - p.hz[o - p.pitch]) - (hy0 - p.hy[o - 1]);
      const float cy = (hx0 - p.hx[o - 1]) - (hz0 - p.hz[o - p.sx]);
      const float cz = (hy0 - p.hy[o - p.sx]) - (hx0 - p.hx[o - p.pitch]);
      const float exn = ade_update<POLES>(p, 0, o, p.ex[o], ca, cb, cx, dk1, dk2, lk1, lk2, lk3);
      const float eyn = ade_update<POLES>(p, 1, o, p.ey[o], ca, cb, cy, dk1, dk2, lk1, lk2, lk3);
      ezn = ade_update<POLES>(p, 2, o, ezn, ca, cb, cz, dk1, dk2, lk1, lk2, lk3);
      p.ex[o] = exn;
      p.ey[o] = eyn;
      if (cell == p.src_cell[bt]) ezn += p.wave[(long long)bt * p.wave_stride_b + p.step];
      p.ez[o] = ezn;
      if (p.e_save) {
        p.e_save[o] = exn;
        p.e_save[p.e_cstride + o] = eyn;
        p.e_save[2 * p.e_cstride + o] = ezn;
      }
    }
    for (int q = probe0; q >= 0; q = p.probe_next[q])
      p.traces[((long long)bt * p.P + q) * p.T + p.step] = ezn;
  }
}

static dim3 grid3(const Step3DArgs& a, int bb) {
  return dim3((a.Mz + 63) / 64, (a.My + 3) / 4, a.Mx * ((a.B + bb - 1) / bb));
}

hipError_t launch_h3d(const Step3DArgs& a, int bb, hipStream_t s) {
  dim3 blk(64, 4, 1), grd = grid3(a, bb);
  const bool mag = a.da != nullptr;
#define CEEP_H3D(M, BBV) hipLaunchKernelGGL((h3d_kernel<M, BBV>), grd, blk, 0, s, a)
  switch (bb) {
    case 1: if (mag) CEEP_H3D(true, 1); else CEEP_H3D(false, 1); break;
    case 2: if (mag) CEEP_H3D(true, 2); else CEEP_H3D(false, 2); break;
    case 4: if (mag) CEEP_H3D(true, 4); else CEEP_H3D(false, 4); break;
    default: return hipErrorInvalidValue;
  }
#undef CEEP_H3D
  return hipGetLastError();
}

template <int POLES>
static hipError_t launch_e3d_p(const Step3DArgs& a, int bb, hipStream_t s) {
  dim3 blk(64, 4, 1), grd = grid3(a, bb);
  switch (bb) {
    case 1: hipLaunchKernelGGL((e3d_kernel<POLES, 1>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((e3d_kernel<POLES, 2>), grd, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((e3d_kernel<POLES, 4>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_e3d(const Step3DArgs& a, int poles, int bb, hipStream_t s) {
  switch (poles) {
    case 0: return launch_e3d_p<0>(a, bb, s);
    case 1: return launch_e3d_p<1>(a, bb, s);
    case 2: return launch_e3d_p<2>(a, bb, s);
    case 3: return launch_e3d_p<3>(a, bb, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ceep
