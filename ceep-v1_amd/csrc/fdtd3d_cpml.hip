// Physical-CPML mode of the batched 3D engine (gfx950).
//
// The batched solvers' own CPML is inert (a PEC cavity); this opt-in mode
// steps the full grid with the reference's convolutional PML instead, in the
// step order of its single-run solver (src/ceep/solvers/fdtd_base.py:161-185):
//   H update on [:, :-1, :-1], [:-1, :, :-1], [:-1, :-1, :]  (fdtd_3d.py:118-133)
//   H psi recursions on the slabs, forward differences     (absorbing.py:627-782)
//   E update on [1:, 1:, 1:] (+ Debye / Lorentz ADE)          (fdtd_3d.py:154-171)
//   Ez soft sources                                           (fdtd_3d.py:173-195)
//   E psi recursions on the slabs, backward differences     (absorbing.py:784-943)
//   Ez probes
// A cell's H corrections read only E and its E corrections only H, so each
// phase is one launch with every correction fused into the cell's update.
// Slab depths: H low k = t-1-j for j in [0, t), H high k = j - (N-1-t) for
// j in [N-1-t, N-1); E low k = t-1-j for j in [1, t), E high k = j - (N-t)
// for j in [N-t, N).  Profiles (b, c) per depth come from the host
// (absorbing.py:266-311).
#include "fdtd3d_kernels.h"

#include <algorithm>

namespace ceep {
namespace {

__device__ __forceinline__ int h_depth(int j, int n, int t) {
  if (j < t) return t - 1 - j;
  if (j >= n - 1 - t && j < n - 1) return j - (n - 1 - t);
  return -1;
}
__device__ __forceinline__ int e_depth(int j, int n, int t) {
  if (j >= 1 && j < t) return t - 1 - j;
  if (j >= n - t && j < n) return j - (n - t);
  return -1;
}

}  // namespace

__global__ __launch_bounds__(256) void cpml3d_h_kernel(Cpml3DArgs p) {
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z % p.Mx;
  const int chunk = blockIdx.z / p.Mx;          // one TX per chunk
  if (b >= p.My || d >= p.Mz || chunk >= p.B) return;
  const long long py = p.pitch, sx = p.sx;
  const long long cell = (long long)a * sx + (long long)b * py + d;
  const long long o = (long long)chunk * p.bstride + cell;
  const bool xa = a < p.Mx - 1, yb = b < p.My - 1, zd = d < p.Mz - 1;
  const int kx = h_depth(a, p.Mx, p.t), ky = h_depth(b, p.My, p.t), kz = h_depth(d, p.Mz, p.t);
  float da = 1.f, c = p.ch;
  if (p.da) { da = p.da[cell]; c = p.dbi[cell]; }
  const float ex0 = p.ex[o], ey0 = p.ey[o], ez0 = p.ez[o];
  const float idx = p.idx, dtmu = p.dtmu;
  if (yb && zd) {   // Hx: dEz/dy (y slabs, -), dEy/dz (z slabs, +)
    const float dzy = p.ez[o + py] - ez0, dyz = p.ey[o + 1] - ey0;
    float h = da * p.hx[o] - c * (dzy - dyz);
    if (ky >= 0) {
      const float ps = p.bh[ky] * p.psi_hxy[o] + p.chh[ky] * dzy * idx;
      p.psi_hxy[o] = ps;
      h -= dtmu * ps;
    }
    if (kz >= 0) {
      const float ps = p.bh[kz] * p.psi_hxz[o] + p.chh[kz] * dyz * idx;
      p.psi_hxz[o] = ps;
      h += dtmu * ps;
    }
    p.hx[o] = h;
  }
  if (xa && zd) {   // Hy: dEx/dz (z slabs, -), dEz/dx (x slabs, +)
    const float dxz = p.ex[o + 1] - ex0, dzx = p.ez[o + sx] - ez0;
    float h = da * p.hy[o] - c * (dxz - dzx);
    if (kz >= 0) {
      const float ps = p.bh[kz] * p.psi_hyz[o] + p.chh[kz] * dxz * idx;
      p.psi_hyz[o] = ps;
      h -= dtmu * ps;
    }
    if (kx >= 0) {
      const float ps = p.bh[kx] * p.psi_hyx[o] + p.chh[kx] * dzx * idx;
      p.psi_hyx[o] = ps;
      h += dtmu * ps;
    }
    p.hy[o] = h;
  }
  if (xa && yb) {   // Hz: dEy/dx (x slabs, -), dEx/dy (y slabs, +)
    const float dyx = p.ey[o + sx] - ey0, dxy = p.ex[o + py] - ex0;
    float h = da * p.hz[o] - c * (dyx - dxy);
    if (kx >= 0) {
      const float ps = p.bh[kx] * p.psi_hzx[o] + p.chh[kx] * dyx * idx;
      p.psi_hzx[o] = ps;
      h -= dtmu * ps;
    }
    if (ky >= 0) {
      const float ps = p.bh[ky] * p.psi_hzy[o] + p.chh[ky] * dxy * idx;
      p.psi_hzy[o] = ps;
      h += dtmu * ps;
    }
    p.hz[o] = h;
  }
}

template <int POLES>
__device__ __forceinline__ float e_update(const Cpml3DArgs& p, int comp, long long o, float e0,
                                          float ca, float cb, float curl, float dk1, float dk2,
                                          float lk1, float lk2, float lk3) {
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

template <int POLES>
__global__ __launch_bounds__(256) void cpml3d_e_kernel(Cpml3DArgs p) {
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z % p.Mx;
  const int bt = blockIdx.z / p.Mx;
  if (b >= p.My || d >= p.Mz || bt >= p.B) return;
  if (a < 1 || b < 1 || d < 1) return;          // the E update domain [1:, 1:, 1:]
  const long long py = p.pitch, sx = p.sx;
  const long long cell = (long long)a * sx + (long long)b * py + d;
  const long long o = (long long)bt * p.bstride + cell;
  const int kx = e_depth(a, p.Mx, p.t), ky = e_depth(b, p.My, p.t), kz = e_depth(d, p.Mz, p.t);
  const float ca = p.ca[cell], cb = p.cb[cell];
  float dk1 = 0.f, dk2 = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
  if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; }
  if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
  const float hx0 = p.hx[o], hy0 = p.hy[o], hz0 = p.hz[o];
  const float dzy = hz0 - p.hz[o - py], dyz = hy0 - p.hy[o - 1];     // Ex
  const float dxz = hx0 - p.hx[o - 1], dzx = hz0 - p.hz[o - sx];     // Ey
  const float dyx = hy0 - p.hy[o - sx], dxy = hx0 - p.hx[o - py];    // Ez
  float ex = e_update<POLES>(p, 0, o, p.ex[o], ca, cb, dzy - dyz, dk1, dk2, lk1, lk2, lk3);
  float ey = e_update<POLES>(p, 1, o, p.ey[o], ca, cb, dxz - dzx, dk1, dk2, lk1, lk2, lk3);
  float ez = e_update<POLES>(p, 2, o, p.ez[o], ca, cb, dyx - dxy, dk1, dk2, lk1, lk2, lk3);
  if (cell == p.src_cell[bt]) ez += p.wave[(long long)bt * p.wave_stride_b + p.step];
  const float idx = p.idx, dte = p.dteps;
  if (ky >= 0) {   // Ex, y slabs (+)
    const float ps = p.be[ky] * p.psi_exy[o] + p.ce[ky] * dzy * idx;
    p.psi_exy[o] = ps;
    ex += dte * ps;
  }
  if (kz >= 0) {   // Ex, z slabs (-)
    const float ps = p.be[kz] * p.psi_exz[o] + p.ce[kz] * dyz * idx;
    p.psi_exz[o] = ps;
    ex -= dte * ps;
  }
  if (kz >= 0) {   // Ey, z slabs (+)
    const float ps = p.be[kz] * p.psi_eyz[o] + p.ce[kz] * dxz * idx;
    p.psi_eyz[o] = ps;
    ey += dte * ps;
  }
  if (kx >= 0) {   // Ey, x slabs (-)
    const float ps = p.be[kx] * p.psi_eyx[o] + p.ce[kx] * dzx * idx;
    p.psi_eyx[o] = ps;
    ey -= dte * ps;
  }
  if (kx >= 0) {   // Ez, x slabs (+)
    const float ps = p.be[kx] * p.psi_ezx[o] + p.ce[kx] * dyx * idx;
    p.psi_ezx[o] = ps;
    ez += dte * ps;
  }
  if (ky >= 0) {   // Ez, y slabs (-)
    const float ps = p.be[ky] * p.psi_ezy[o] + p.ce[ky] * dxy * idx;
    p.psi_ezy[o] = ps;
    ez -= dte * ps;
  }
  p.ex[o] = ex;
  p.ey[o] = ey;
  p.ez[o] = ez;
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
  for (int q = probe0; q >= 0; q = p.probe_next[q])
    p.traces[((long long)bt * p.P + q) * p.T + p.step] = ez;
}

hipError_t launch_cpml3d(const Cpml3DArgs& a, int poles, hipStream_t s) {
  if (a.t < 1 || a.t > kMaxCpml || 2 * a.t + 1 > a.Mx || 2 * a.t + 1 > a.My || 2 * a.t + 1 > a.Mz)
    return hipErrorInvalidValue;
  const long long gz = (long long)a.Mx * a.B;
  if (gz > 65535) return hipErrorInvalidValue;   // grid z limit
  const dim3 blk(64, 4, 1);
  const dim3 grd((a.Mz + 63) / 64, (a.My + 3) / 4, (unsigned)gz);
  hipLaunchKernelGGL(cpml3d_h_kernel, grd, blk, 0, s, a);
  switch (poles) {
    case 0: hipLaunchKernelGGL((cpml3d_e_kernel<0>), grd, blk, 0, s, a); break;
    case 1: hipLaunchKernelGGL((cpml3d_e_kernel<1>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((cpml3d_e_kernel<2>), grd, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((cpml3d_e_kernel<3>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ceep
