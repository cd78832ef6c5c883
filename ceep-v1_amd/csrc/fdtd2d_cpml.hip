// Physical-CPML mode of the batched 2D TMz engine (gfx950).
//
// The batched solvers' own CPML is inert (a PEC cavity); this opt-in mode
// steps the full grid with the reference's convolutional PML instead, in the
// step order of its single-run solver (src/ceep/solvers/fdtd_base.py:161-185):
//   H update on [:, :-1] / [:-1, :]           (fdtd_2d.py:129-140)
//   H psi recursions on the four slabs         (absorbing.py:405-453)
//   E update on [1:, 1:] (+ Debye / Lorentz ADE) (fdtd_2d.py:163-178)
//   soft sources                               (fdtd_2d.py:237-246)
//   E psi recursions on the four slabs         (absorbing.py:513-561)
//   probes
// Profiles (b, c) per depth come from the host (absorbing.py:266-311).  Two
// launches per step; fields in place.  psi lives in four batch-plane arrays
// (x-slab rows and y-slab columns of one kind share an array: the low and
// high slabs never overlap).
#include "fdtd2d_kernels.h"

#include <algorithm>

namespace ceep {

__global__ __launch_bounds__(256) void cpml2d_h_kernel(Cpml2DArgs p) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  const int a = blockIdx.y * 4 + threadIdx.y;
  if (a >= p.Mx || c >= p.My) return;
  const long long cell = (long long)a * p.pitch + c;
  const int t = p.t;
  const bool has_r = c < p.My - 1, has_d = a < p.Mx - 1;
  // depth indices into the profiles (-1: not in that slab)
  const int xlo = (has_d && a < t) ? t - 1 - a : -1;
  const int xhi = (has_d && a >= p.Mx - t) ? a - (p.Mx - t) : -1;
  const int ylo = (has_r && c < t) ? t - 1 - c : -1;
  const int yhi = (has_r && c >= p.My - t) ? c - (p.My - t) : -1;
  for (int bt = blockIdx.z; bt < p.B; bt += gridDim.z) {
    const long long o = (long long)bt * p.bstride + cell;
    const float e0 = p.ez[o];
    if (has_r) {
      const float de = p.ez[o + 1] - e0;
      float h = p.hx[o] - p.dbdy * de;
      if (ylo >= 0 || yhi >= 0) {
        const int k = ylo >= 0 ? ylo : yhi;
        const float ps = p.bh[k] * p.psi_hxy[o] + p.chh[k] * de * p.idx;
        p.psi_hxy[o] = ps;
        h -= p.dtmu * ps;
      }
      p.hx[o] = h;
    }
    if (has_d) {
      const float de = p.ez[o + p.pitch] - e0;
      float h = p.hy[o] + p.dbdy * de;
      if (xlo >= 0 || xhi >= 0) {
        const int k = xlo >= 0 ? xlo : xhi;
        const float ps = p.bh[k] * p.psi_hyx[o] + p.chh[k] * de * p.idx;
        p.psi_hyx[o] = ps;
        h += p.dtmu * ps;
      }
      p.hy[o] = h;
    }
  }
}

template <int POLES>
__global__ __launch_bounds__(256) void cpml2d_e_kernel(Cpml2DArgs p) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  const int a = blockIdx.y * 4 + threadIdx.y;
  if (a >= p.Mx || c >= p.My) return;
  const long long cell = (long long)a * p.pitch + c;
  const int t = p.t;
  const bool main = a >= 1 && c >= 1;   // the E update domain [1:, 1:]
  const int xlo = (a >= 1 && a < t) ? t - 1 - a : -1;
  const int xhi = (a >= 1 && a >= p.Mx - t) ? a - (p.Mx - t) : -1;
  const int ylo = (c >= 1 && c < t) ? t - 1 - c : -1;
  const int yhi = (c >= 1 && c >= p.My - t) ? c - (p.My - t) : -1;
  float ca = 0.f, cb = 0.f, dk1 = 0.f, dk2 = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
  if (main) {
    ca = p.ca[cell];
    cb = p.cb[cell];
    if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; }
    if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
  }
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
  for (int bt = blockIdx.z; bt < p.B; bt += gridDim.z) {
    const long long o = (long long)bt * p.bstride + cell;
    float e = p.ez[o];
    const float dhy = (a >= 1) ? p.hy[o] - p.hy[o - p.pitch] : 0.f;
    const float dhx = (c >= 1) ? p.hx[o] - p.hx[o - 1] : 0.f;
    if (main) {
      const float curl = dhy * p.idx - dhx * p.idx;
      if (POLES == 0) {
        e = ca * e + cb * curl;
      } else {
        float jd = 0.f, jl = 0.f, jlp = 0.f, sgj = 0.f;
        if (POLES & 1) { jd = p.jd[o]; sgj += 0.5f * (1.f + dk1) * jd; }
        if (POLES & 2) { jl = p.jl_cur[o]; jlp = p.jl_prev[o]; sgj += jl; }
        const float en = ca * e + cb * curl - cb * sgj;
        const float de = en - e;
        if (POLES & 1) p.jd[o] = dk1 * jd + dk2 * de;
        if (POLES & 2) p.jl_prev[o] = lk1 * jl + lk3 * jlp + lk2 * de;
        e = en;
      }
    }
    if (cell == p.src_cell[bt]) e += p.wave[(long long)bt * p.wave_stride_b + p.step];
    if (xlo >= 0 || xhi >= 0) {
      const int k = xlo >= 0 ? xlo : xhi;
      const float ps = p.be[k] * p.psi_ezx[o] + p.ce[k] * dhy * p.idx;
      p.psi_ezx[o] = ps;
      e += p.dteps * ps;
    }
    if (ylo >= 0 || yhi >= 0) {
      const int k = ylo >= 0 ? ylo : yhi;
      const float ps = p.be[k] * p.psi_ezy[o] + p.ce[k] * dhx * p.idx;
      p.psi_ezy[o] = ps;
      e -= p.dteps * ps;
    }
    p.ez[o] = e;
    for (int q = probe0; q >= 0; q = p.probe_next[q])
      p.traces[((long long)bt * p.P + q) * p.T + p.step] = e;
  }
}

hipError_t launch_cpml2d(const Cpml2DArgs& a, int poles, hipStream_t s) {
  if (a.t < 1 || a.t > kMaxCpml || 2 * a.t > a.Mx || 2 * a.t > a.My) return hipErrorInvalidValue;
  const dim3 blk(64, 4, 1);
  const dim3 grd((a.My + 63) / 64, (a.Mx + 3) / 4, (unsigned)std::min(a.B, 16));
  hipLaunchKernelGGL(cpml2d_h_kernel, grd, blk, 0, s, a);
  switch (poles) {
    case 0: hipLaunchKernelGGL((cpml2d_e_kernel<0>), grd, blk, 0, s, a); break;
    case 1: hipLaunchKernelGGL((cpml2d_e_kernel<1>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((cpml2d_e_kernel<2>), grd, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((cpml2d_e_kernel<3>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ceep
