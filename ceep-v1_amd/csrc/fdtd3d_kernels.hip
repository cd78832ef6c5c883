// Batched 3D Yee step kernels for gfx950.
//
// H kernel: the three H components in place (fdtd_3d_batched.py:753-818, on
// the cavity box, so the reference's H-CPML is the identity there).
// E kernel: the three E components in place on the E interior
// (fdtd_3d_batched.py:820-884) with the ADE of fdtd_2d.py:163-178 per
// component, then Ez injection (:949-951) and Ez probes (:956-958).
//
// Mapping: a 256-thread workgroup covers 64 consecutive z-cells (the
// contiguous axis) x 4 y-rows of one x-slab for BB batch planes, so the
// material maps are read once per chunk.  The 1D grid enumerates (z/y tile,
// x-slab, batch chunk) with the tile fastest; with xcd_order the blocks are
// renumbered so each XCD walks a contiguous run of slabs and the +-1-slab
// neighbour reads hit its L2.  Every thread issues all loads of its BB
// planes before it computes or stores anything: these kernels are bound by
// HBM, and the loads in flight per wave are what feeds it.  Floating-point
// contraction is off, so the other 3D step forms reproduce these results bit
// for bit from the same expressions.
#include "fdtd_common.h"
#include "fdtd3d_kernels.h"

namespace ceep {

namespace {

struct Block3D {
  int a, b, d, chunk;
};

// xcd_order 0: dispatch order, tile fastest, then slab, then batch chunk;
// 1: the same order renumbered per XCD; 2: per XCD with the batch chunk
// fastest, so the chunks of one tile run together and share its material
// maps through that XCD's L2.
__device__ __forceinline__ Block3D block3d(const Step3DArgs& p, int bb) {
  const int nbz = (p.Mz + 63) / 64, nby = (p.My + 3) / 4;
  const int tiles = nbz * nby;
  const int L = xcd_block(p.xcd_order);
  Block3D r;
  int tile;
  if (p.xcd_order == 2) {
    const int nch = (p.B + bb - 1) / bb;
    r.chunk = L % nch;
    const int rest = L / nch;
    tile = rest % tiles;
    r.a = rest / tiles;
  } else {
    tile = L % tiles;
    const int rest = L / tiles;
    r.a = rest % p.Mx;
    r.chunk = rest / p.Mx;
  }
  r.d = (tile % nbz) * 64 + (threadIdx.x & 63);
  r.b = (tile / nbz) * 4 + (threadIdx.x >> 6);
  return r;
}

}  // namespace

template <bool MAGNETIC, int BB>
__global__ __launch_bounds__(256) void h3d_kernel(Step3DArgs p) {
#pragma clang fp contract(off)
  const Block3D w = block3d(p, BB);
  const int a = w.a, b = w.b, d = w.d;
  if (b >= p.My || d >= p.Mz) return;
  const long long cell = (long long)a * p.sx + (long long)b * p.pitch + d;
  const long long py = p.pitch, sx = p.sx;
  const bool xa = a < p.Mx - 1, yb = b < p.My - 1, zd = d < p.Mz - 1;
  float da = 1.f, c = p.ch;
  if (MAGNETIC) { da = p.da[cell]; c = p.dbi[cell]; }
  const int b0 = w.chunk * BB;
  const int nb = min(BB, p.B - b0);
  const float* __restrict__ ex = p.ex;
  const float* __restrict__ ey = p.ey;
  const float* __restrict__ ez = p.ez;
  // all loads first
  float e0[BB][3], eyz[BB], ezy[BB], ezx[BB], exz[BB], exy[BB], eyx[BB], h0[BB][3];
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    if (k >= nb) break;
    const long long o = (long long)(b0 + k) * p.bstride + cell;
    e0[k][0] = ex[o]; e0[k][1] = ey[o]; e0[k][2] = ez[o];
    // hx: y < My-1, z < Mz-1 ; hy: x < Mx-1, z < Mz-1 ; hz: x < Mx-1, y < My-1
    if (yb && zd) { eyz[k] = ey[o + 1]; ezy[k] = ez[o + py]; h0[k][0] = p.hx[o]; }
    if (xa && zd) { ezx[k] = ez[o + sx]; exz[k] = ex[o + 1]; h0[k][1] = p.hy[o]; }
    if (xa && yb) { exy[k] = ex[o + py]; eyx[k] = ey[o + sx]; h0[k][2] = p.hz[o]; }
  }
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    if (k >= nb) break;
    const long long o = (long long)(b0 + k) * p.bstride + cell;
    if (yb && zd) {
      const float t = (eyz[k] - e0[k][1]) - (ezy[k] - e0[k][2]);
      p.hx[o] = MAGNETIC ? da * h0[k][0] + c * t : h0[k][0] + c * t;
    }
    if (xa && zd) {
      const float t = (ezx[k] - e0[k][2]) - (exz[k] - e0[k][0]);
      p.hy[o] = MAGNETIC ? da * h0[k][1] + c * t : h0[k][1] + c * t;
    }
    if (xa && yb) {
      const float t = (exy[k] - e0[k][0]) - (eyx[k] - e0[k][1]);
      p.hz[o] = MAGNETIC ? da * h0[k][2] + c * t : h0[k][2] + c * t;
    }
  }
}

template <int POLES, int BB>
__global__ __launch_bounds__(256) void e3d_kernel(Step3DArgs p) {
#pragma clang fp contract(off)
  const Block3D w = block3d(p, BB);
  const int a = w.a, b = w.b, d = w.d;
  if (b >= p.My || d >= p.Mz) return;
  const long long cell = (long long)a * p.sx + (long long)b * p.pitch + d;
  const long long py = p.pitch, sx = p.sx;
  const bool interior = a >= 1 && a <= p.Mx - 2 && b >= 1 && b <= p.My - 2 &&
                        d >= 1 && d <= p.Mz - 2;
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
  if (!interior && probe0 < 0) return;
  const int b0 = w.chunk * BB;
  const int nb = min(BB, p.B - b0);
  if (!interior) {   // a probe off the E interior records the unchanged Ez
#pragma unroll
    for (int k = 0; k < BB; ++k) {
      if (k >= nb) break;
      const int bt = b0 + k;
      const float v = p.ez[(long long)bt * p.bstride + cell];
      for (int q = probe0; q >= 0; q = p.probe_next[q])
        p.traces[((long long)bt * p.P + q) * p.T + p.step] = v;
    }
    return;
  }
  const float ca = p.ca[cell], cb = p.cb[cell];
  float dk1 = 0.f, dk2 = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
  if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; }
  if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
  const float* __restrict__ hx = p.hx;
  const float* __restrict__ hy = p.hy;
  const float* __restrict__ hz = p.hz;
  // all loads first: E, H and its -1 neighbours, the ADE currents
  float e0[BB][3], cu[BB][3], jd[BB][3], jl[BB][3], jlp[BB][3];
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    if (k >= nb) break;
    const long long o = (long long)(b0 + k) * p.bstride + cell;
    e0[k][0] = p.ex[o]; e0[k][1] = p.ey[o]; e0[k][2] = p.ez[o];
    const float hx0 = hx[o], hy0 = hy[o], hz0 = hz[o];
    cu[k][0] = (hz0 - hz[o - py]) - (hy0 - hy[o - 1]);
    cu[k][1] = (hx0 - hx[o - 1]) - (hz0 - hz[o - sx]);
    cu[k][2] = (hy0 - hy[o - sx]) - (hx0 - hx[o - py]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (POLES & 1) jd[k][c] = p.jd[c][o];
      if (POLES & 2) { jl[k][c] = p.jl_cur[c][o]; jlp[k][c] = p.jl_prev[c][o]; }
    }
  }
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    if (k >= nb) break;
    const int bt = b0 + k;
    const long long o = (long long)bt * p.bstride + cell;
    float en[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (POLES == 0) {
        en[c] = ca * e0[k][c] + cb * (cu[k][c] * p.idx);
      } else {
        float sgj = 0.f;
        if (POLES & 1) sgj += 0.5f * (1.f + dk1) * jd[k][c];
        if (POLES & 2) sgj += jl[k][c];
        en[c] = ca * e0[k][c] + cb * (cu[k][c] * p.idx) - cb * sgj;
        const float de = en[c] - e0[k][c];
        if (POLES & 1) p.jd[c][o] = dk1 * jd[k][c] + dk2 * de;
        if (POLES & 2) p.jl_prev[c][o] = lk1 * jl[k][c] + lk3 * jlp[k][c] + lk2 * de;
      }
    }
    p.ex[o] = en[0];
    p.ey[o] = en[1];
    float ezn = en[2];
    if (cell == p.src_cell[bt]) ezn += p.wave[(long long)bt * p.wave_stride_b + p.step];
    p.ez[o] = ezn;
    if (p.e_save) {
      p.e_save[o] = en[0];
      p.e_save[p.e_cstride + o] = en[1];
      p.e_save[2 * p.e_cstride + o] = ezn;
    }
    for (int q = probe0; q >= 0; q = p.probe_next[q])
      p.traces[((long long)bt * p.P + q) * p.T + p.step] = ezn;
  }
}

static dim3 grid3(const Step3DArgs& a, int bb) {
  const long long n = (long long)((a.Mz + 63) / 64) * ((a.My + 3) / 4) * a.Mx *
                      ((a.B + bb - 1) / bb);
  return dim3((unsigned)n, 1, 1);
}

hipError_t launch_h3d(const Step3DArgs& a, int bb, hipStream_t s) {
  dim3 blk(256, 1, 1), grd = grid3(a, bb);
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
  dim3 blk(256, 1, 1), grd = grid3(a, bb);
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
