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

// every field buffer of a fused launch must be set
static bool fused_buffers_set(const Step3DFArgs& f) {
  for (int k = 0; k < 3; ++k)
    if (!f.e_in[k] || !f.h_in[k] || !f.e_out[k] || !f.h_out[k]) return false;
  return true;
}

hipError_t launch_step3d_fused(const Step3DFArgs& f, int poles, int bb, hipStream_t s) {
  if (!fused_buffers_set(f)) return hipErrorInvalidValue;
  switch (poles) {
    case 0: return launch_fused_p<0>(f, bb, s);
    case 1: return launch_fused_p<1>(f, bb, s);
    case 2: return launch_fused_p<2>(f, bb, s);
    case 3: return launch_fused_p<3>(f, bb, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ceep

namespace ceep {

// Marching form of step3d_fused_kernel.  Workgroup = 64 z-cells x 4 y-rows x
// one TX chunk x one run of seg_len x-slabs; the thread walks its (y, z)
// column from slab a0 to a0 + seg_len - 1.  At slab a it loads E_in(a+1) of
// its own cell (kept as the next slab's E), computes its own new H(a) (hx,
// hy, hz) and, on the E interior, E(a) from
//   new hx, hy at (a, y, z-1) and hx, hz at (a, y-1, z): recomputed from the
//       input fields exactly as their owners compute them (contraction off)
//   new hy, hz at (a-1, y, z): this thread's own values from slab a-1.
// In-slab neighbour loads are L1/L2 hits (neighbouring threads load the same
// slab); each slab of E_in, H_in is fetched from HBM once per step.
template <bool MAGNETIC, int POLES, int BB>
__global__ __launch_bounds__(256) void step3d_march_kernel(Step3DFArgs f) {
#pragma clang fp contract(off)
  const Step3DArgs& p = f.a;
  const int nbz = (p.Mz + 63) / 64, nby = (p.My + 3) / 4;
  const int tiles = nbz * nby;
  const int nseg = (p.Mx + f.seg_len - 1) / f.seg_len;
  const int blk = blockIdx.x;
  const int tile = blk % tiles;
  const int rest = blk / tiles;
  const int seg = rest % nseg;
  const int chunk = rest / nseg;
  const int d = (tile % nbz) * 64 + (threadIdx.x & 63);
  const int b = (tile / nbz) * 4 + (threadIdx.x >> 6);
  if (b >= p.My || d >= p.Mz) return;
  const int a_begin = seg * f.seg_len;
  const int a_end = min(a_begin + f.seg_len, p.Mx);
  const long long sx = p.sx, py = p.pitch;
  const long long col = (long long)b * py + d;          // cell of slab 0
  const bool yb = b < p.My - 1, zd = d < p.Mz - 1;
  const bool yz_in = b >= 1 && yb && d >= 1 && zd;      // interior in y and z
  const float* __restrict__ ex = f.e_in[0];
  const float* __restrict__ ey = f.e_in[1];
  const float* __restrict__ ez = f.e_in[2];
  const float* __restrict__ hx = f.h_in[0];
  const float* __restrict__ hy = f.h_in[1];
  const float* __restrict__ hz = f.h_in[2];
  const int b0 = chunk * BB;
  const int nb = min(BB, p.B - b0);

  // carried per TX plane: E_in of the current slab, new hy/hz of the previous slab
  float ecx[BB], ecy[BB], ecz[BB], hpy[BB], hpz[BB];
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    ecx[k] = ecy[k] = ecz[k] = hpy[k] = hpz[k] = 0.f;
    if (k >= nb) continue;
    const long long o = (long long)(b0 + k) * p.bstride + (long long)a_begin * sx + col;
    ecx[k] = ex[o];
    ecy[k] = ey[o];
    ecz[k] = ez[o];
    if (a_begin > 0) {
      // own new hy, hz at slab a_begin - 1 (the previous run writes them)
      const long long om = o - sx;
      const float c = MAGNETIC ? p.dbi[(long long)(a_begin - 1) * sx + col] : p.ch;
      const float dd = MAGNETIC ? p.da[(long long)(a_begin - 1) * sx + col] : 1.f;
      const float exm = ex[om], eym = ey[om], ezm = ez[om];
      if (zd) {
        const float t = (ecz[k] - ezm) - (ex[om + 1] - exm);
        hpy[k] = MAGNETIC ? dd * hy[om] + c * t : hy[om] + c * t;
      }
      if (yb) {
        const float t = (ex[om + py] - exm) - (ecy[k] - eym);
        hpz[k] = MAGNETIC ? dd * hz[om] + c * t : hz[om] + c * t;
      }
    }
  }

  for (int a = a_begin; a < a_end; ++a) {
    const long long cell = (long long)a * sx + col;
    const bool xa = a < p.Mx - 1;
    const bool interior = yz_in && a >= 1 && xa;
    const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
    float da0 = 1.f, c0 = p.ch, dmz = 1.f, cmz = p.ch, dmy = 1.f, cmy = p.ch;
    if (MAGNETIC) {
      da0 = p.da[cell]; c0 = p.dbi[cell];
    }
    float ca = 0.f, cb = 0.f, dk1 = 0.f, dk2 = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
    if (interior) {
      if (MAGNETIC) {
        dmz = p.da[cell - 1]; cmz = p.dbi[cell - 1];
        dmy = p.da[cell - py]; cmy = p.dbi[cell - py];
      }
      ca = p.ca[cell];
      cb = p.cb[cell];
      if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; }
      if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
    }
#pragma unroll
    for (int k = 0; k < BB; ++k) {
      if (k >= nb) break;
      const int bt = b0 + k;
      const long long o = (long long)bt * p.bstride + cell;
      const float ex0 = ecx[k], ey0 = ecy[k], ez0 = ecz[k];
      float enx = 0.f, eny = 0.f, enz = 0.f;   // E_in of slab a+1 (this cell)
      if (xa) {
        enx = ex[o + sx];
        eny = ey[o + sx];
        enz = ez[o + sx];
      }
      float hx0 = 0.f, hy0 = 0.f, hz0 = 0.f;
      if (yb && zd) {
        const float t = (ey[o + 1] - ey0) - (ez[o + py] - ez0);
        hx0 = MAGNETIC ? da0 * hx[o] + c0 * t : hx[o] + c0 * t;
        f.h_out[0][o] = hx0;
      }
      if (xa && zd) {
        const float t = (enz - ez0) - (ex[o + 1] - ex0);
        hy0 = MAGNETIC ? da0 * hy[o] + c0 * t : hy[o] + c0 * t;
        f.h_out[1][o] = hy0;
      }
      if (xa && yb) {
        const float t = (ex[o + py] - ex0) - (eny - ey0);
        hz0 = MAGNETIC ? da0 * hz[o] + c0 * t : hz[o] + c0 * t;
        f.h_out[2][o] = hz0;
      }
      float ezn = ez0;
      if (interior) {
        const long long oz = o - 1, oy = o - py;
        const float ey_z = ey[oz], ez_z = ez[oz], ex_z = ex[oz];
        const float ez_y = ez[oy], ex_y = ex[oy], ey_y = ey[oy];
        float t;
        t = (ey0 - ey_z) - (ez[oz + py] - ez_z);                 // hx(a, y, z-1)
        const float hx_mz = MAGNETIC ? dmz * hx[oz] + cmz * t : hx[oz] + cmz * t;
        t = (ez[oz + sx] - ez_z) - (ex0 - ex_z);                 // hy(a, y, z-1)
        const float hy_mz = MAGNETIC ? dmz * hy[oz] + cmz * t : hy[oz] + cmz * t;
        t = (ey[oy + 1] - ey_y) - (ez0 - ez_y);                  // hx(a, y-1, z)
        const float hx_my = MAGNETIC ? dmy * hx[oy] + cmy * t : hx[oy] + cmy * t;
        t = (ex0 - ex_y) - (ey[oy + sx] - ey_y);                 // hz(a, y-1, z)
        const float hz_my = MAGNETIC ? dmy * hz[oy] + cmy * t : hz[oy] + cmy * t;
        const float cx = (hz0 - hz_my) - (hy0 - hy_mz);
        const float cy = (hx0 - hx_mz) - (hz0 - hpz[k]);
        const float cz = (hy0 - hpy[k]) - (hx0 - hx_my);
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
      ecx[k] = enx;
      ecy[k] = eny;
      ecz[k] = enz;
      hpy[k] = hy0;
      hpz[k] = hz0;
    }
  }
}

template <bool MAGNETIC, int POLES>
static hipError_t launch_march_mp(const Step3DFArgs& f, int bb, hipStream_t s) {
  const Step3DArgs& a = f.a;
  if (f.seg_len < 1) return hipErrorInvalidValue;
  const long long nseg = (a.Mx + f.seg_len - 1) / f.seg_len;
  const long long blocks =
      (long long)((a.Mz + 63) / 64) * ((a.My + 3) / 4) * nseg * ((a.B + bb - 1) / bb);
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 blk(256, 1, 1), grd((unsigned)blocks, 1, 1);
  switch (bb) {
    case 1: hipLaunchKernelGGL((step3d_march_kernel<MAGNETIC, POLES, 1>), grd, blk, 0, s, f); break;
    case 2: hipLaunchKernelGGL((step3d_march_kernel<MAGNETIC, POLES, 2>), grd, blk, 0, s, f); break;
    case 4: hipLaunchKernelGGL((step3d_march_kernel<MAGNETIC, POLES, 4>), grd, blk, 0, s, f); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int POLES>
static hipError_t launch_march_p(const Step3DFArgs& f, int bb, hipStream_t s) {
  return f.a.da ? launch_march_mp<true, POLES>(f, bb, s) : launch_march_mp<false, POLES>(f, bb, s);
}

hipError_t launch_step3d_march(const Step3DFArgs& f, int poles, int bb, hipStream_t s) {
  if (!fused_buffers_set(f)) return hipErrorInvalidValue;
  switch (poles) {
    case 0: return launch_march_p<0>(f, bb, s);
    case 1: return launch_march_p<1>(f, bb, s);
    case 2: return launch_march_p<2>(f, bb, s);
    case 3: return launch_march_p<3>(f, bb, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ceep
