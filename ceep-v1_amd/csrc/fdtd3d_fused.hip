// Fused batched 3D Yee step for gfx950: H half step and E step (+ Debye /
// Lorentz ADE, Ez injection, Ez probes) in one launch, ping-pong buffers.
//
// Same update as h3d_kernel + e3d_kernel (fdtd_3d_batched.py:753-884,
// 949-958; ADE of fdtd_2d.py:163-178), but every E cell recomputes the six
// new H values its curl needs at the -1 neighbours from the input fields
// (the thread that owns such a cell computes the identical value; floating-
// point contraction is off so both agree bit for bit).  The step therefore
// reads E and H once and writes them once: 48 B of field traffic per
// cell-TX instead of 72 B for the two-kernel form.
//
// Mapping: 256 threads = 64 z-cells x 4 y-rows of one x-slab; BB TX planes
// per thread share the material loads.  With xcd_order the workgroups are
// renumbered so each of the 8 XCDs (which receive workgroups round-robin)
// walks a contiguous run of x-slabs, keeping the slab a-1 data an E update
// re-reads in that XCD's L2.
#include "fdtd3d_kernels.h"

namespace ceep {

template <bool MAGNETIC, int POLES, int BB>
__global__ __launch_bounds__(256) void step3d_fused_kernel(Step3DFArgs f) {
#pragma clang fp contract(off)
  const Step3DArgs& p = f.a;
  const int nbz = (p.Mz + 63) / 64, nby = (p.My + 3) / 4;
  const int tiles = nbz * nby;
  // XCD-contiguous renumbering of the 1D grid
  const int N = gridDim.x;
  const int phys = blockIdx.x;
  const int q = N / 8, r = N % 8, xcd = phys % 8;
  const int logical = f.xcd_order ? xcd * q + min(xcd, r) + phys / 8 : phys;
  const int chunk = logical / (tiles * p.Mx);
  const int rem = logical - chunk * tiles * p.Mx;
  const int a = rem / tiles;
  const int tile = rem - a * tiles;
  const int d = (tile % nbz) * 64 + (threadIdx.x & 63);
  const int b = (tile / nbz) * 4 + (threadIdx.x >> 6);
  if (b >= p.My || d >= p.Mz) return;
  const long long sx = p.sx, py = p.pitch;
  const long long cell = (long long)a * sx + (long long)b * py + d;
  const bool xa = a < p.Mx - 1, yb = b < p.My - 1, zd = d < p.Mz - 1;
  const bool interior = a >= 1 && xa && b >= 1 && yb && d >= 1 && zd;
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;

  float da0 = 1.f, c0 = p.ch;
  float cmz = p.ch, cmy = p.ch, cmx = p.ch, dmz = 1.f, dmy = 1.f, dmx = 1.f;
  if (MAGNETIC) {
    da0 = p.da[cell];
    c0 = p.dbi[cell];
  }
  float ca = 0.f, cb = 0.f, dk1 = 0.f, dk2 = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
  if (interior) {
    if (MAGNETIC) {
      dmz = p.da[cell - 1]; cmz = p.dbi[cell - 1];
      dmy = p.da[cell - py]; cmy = p.dbi[cell - py];
      dmx = p.da[cell - sx]; cmx = p.dbi[cell - sx];
    }
    ca = p.ca[cell];
    cb = p.cb[cell];
    if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; }
    if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
  }
  const float* __restrict__ ex = f.e_in[0];
  const float* __restrict__ ey = f.e_in[1];
  const float* __restrict__ ez = f.e_in[2];
  const float* __restrict__ hx = f.h_in[0];
  const float* __restrict__ hy = f.h_in[1];
  const float* __restrict__ hz = f.h_in[2];

  const int b0 = chunk * BB;
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    const int bt = b0 + k;
    if (bt >= p.B) break;
    const long long o = (long long)bt * p.bstride + cell;
    const float ex0 = ex[o], ey0 = ey[o], ez0 = ez[o];
    // own H (update ranges of h3d_kernel)
    float hx0 = 0.f, hy0 = 0.f, hz0 = 0.f;
    if (yb && zd) {
      const float t = (ey[o + 1] - ey0) - (ez[o + py] - ez0);
      hx0 = MAGNETIC ? da0 * hx[o] + c0 * t : hx[o] + c0 * t;
      f.h_out[0][o] = hx0;
    }
    if (xa && zd) {
      const float t = (ez[o + sx] - ez0) - (ex[o + 1] - ex0);
      hy0 = MAGNETIC ? da0 * hy[o] + c0 * t : hy[o] + c0 * t;
      f.h_out[1][o] = hy0;
    }
    if (xa && yb) {
      const float t = (ex[o + py] - ex0) - (ey[o + sx] - ey0);
      hz0 = MAGNETIC ? da0 * hz[o] + c0 * t : hz[o] + c0 * t;
      f.h_out[2][o] = hz0;
    }
    float ezn = ez0;
    if (interior) {
      // new H at the -1 neighbours, exactly as their owners compute them
      const long long oz = o - 1, oy = o - py, ox = o - sx;
      const float ey_z = ey[oz], ez_z = ez[oz], ex_z = ex[oz];
      const float ez_y = ez[oy], ex_y = ex[oy], ey_y = ey[oy];
      const float ex_x = ex[ox], ey_x = ey[ox], ez_x = ez[ox];
      float t;
      t = (ey0 - ey_z) - (ez[oz + py] - ez_z);                 // hx(o - 1)
      const float hx_mz = MAGNETIC ? dmz * hx[oz] + cmz * t : hx[oz] + cmz * t;
      t = (ez[oz + sx] - ez_z) - (ex0 - ex_z);                 // hy(o - 1)
      const float hy_mz = MAGNETIC ? dmz * hy[oz] + cmz * t : hy[oz] + cmz * t;
      t = (ey[oy + 1] - ey_y) - (ez0 - ez_y);                  // hx(o - pitch)
      const float hx_my = MAGNETIC ? dmy * hx[oy] + cmy * t : hx[oy] + cmy * t;
      t = (ex0 - ex_y) - (ey[oy + sx] - ey_y);                 // hz(o - pitch)
      const float hz_my = MAGNETIC ? dmy * hz[oy] + cmy * t : hz[oy] + cmy * t;
      t = (ez0 - ez_x) - (ex[ox + 1] - ex_x);                  // hy(o - sx)
      const float hy_mx = MAGNETIC ? dmx * hy[ox] + cmx * t : hy[ox] + cmx * t;
      t = (ex[ox + py] - ex_x) - (ey0 - ey_x);                 // hz(o - sx)
      const float hz_mx = MAGNETIC ? dmx * hz[ox] + cmx * t : hz[ox] + cmx * t;
      const float cx = (hz0 - hz_my) - (hy0 - hy_mz);
      const float cy = (hx0 - hx_mz) - (hz0 - hz_mx);
      const float cz = (hy0 - hy_mx) - (hx0 - hx_my);
      const float e0[3] = {ex0, ey0, ez0};
      const float cu[3] = {cx, cy, cz};
      float en[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (POLES == 0) {
          en[c] = ca * e0[c] + cb * (cu[c] * p.idx);
        } else {
          float jd = 0.f, jl = 0.f, jlp = 0.f, sgj = 0.f;
          if (POLES & 1) { jd = p.jd[c][o]; sgj += 0.5f * (1.f + dk1) * jd; }
          if (POLES & 2) { jl = p.jl_cur[c][o]; jlp = p.jl_prev[c][o]; sgj += jl; }
          en[c] = ca * e0[c] + cb * (cu[c] * p.idx) - cb * sgj;
          const float de = en[c] - e0[c];
          if (POLES & 1) p.jd[c][o] = dk1 * jd + dk2 * de;
          if (POLES & 2) p.jl_prev[c][o] = lk1 * jl + lk3 * jlp + lk2 * de;
        }
      }
      ezn = en[2];
      if (cell == p.src_cell[bt]) ezn += p.wave[(long long)bt * p.wave_stride_b + p.step];
      f.e_out[0][o] = en[0];
      f.e_out[1][o] = en[1];
      f.e_out[2][o] = ezn;
      if (p.e_save) {
        p.e_save[o] = en[0];
        p.e_save[p.e_cstride + o] = en[1];
        p.e_save[2 * p.e_cstride + o] = ezn;
      }
    }
    for (int pq = probe0; pq >= 0; pq = p.probe_next[pq])
      p.traces[((long long)bt * p.P + pq) * p.T + p.step] = ezn;
  }
}

template <bool MAGNETIC, int POLES>
static hipError_t launch_fused_mp(const Step3DFArgs& f, int bb, hipStream_t s) {
  const Step3DArgs& a = f.a;
  const long long blocks =
      (long long)((a.Mz + 63) / 64) * ((a.My + 3) / 4) * a.Mx * ((a.B + bb - 1) / bb);
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 blk(256, 1, 1), grd((unsigned)blocks, 1, 1);
  switch (bb) {
    case 1: hipLaunchKernelGGL((step3d_fused_kernel<MAGNETIC, POLES, 1>), grd, blk, 0, s, f); break;
    case 2: hipLaunchKernelGGL((step3d_fused_kernel<MAGNETIC, POLES, 2>), grd, blk, 0, s, f); break;
    case 4: hipLaunchKernelGGL((step3d_fused_kernel<MAGNETIC, POLES, 4>), grd, blk, 0, s, f); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int POLES>
static hipError_t launch_fused_p(const Step3DFArgs& f, int bb, hipStream_t s) {
  return f.a.da ? launch_fused_mp<true, POLES>(f, bb, s) : launch_fused_mp<false, POLES>(f, bb, s);
}

hipError_t launch_step3d_fused(const Step3DFArgs& f, int poles, int bb, hipStream_t s) {
  switch (poles) {
    case 0: return launch_fused_p<0>(f, bb, s);
    case 1: return launch_fused_p<1>(f, bb, s);
    case 2: return launch_fused_p<2>(f, bb, s);
    case 3: return launch_fused_p<3>(f, bb, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ceep
