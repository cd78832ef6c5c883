// Time-reversed (discrete adjoint) step of the batched 3D update for gfx950.
//
// The forward step (fdtd3d_kernels.hip) is, per TX and cell,
//   h'  = da h + dbi dE(e)                      H half step, in its update range
//   e~  = ca e + cb/dx curl(h') - cb (gD JD + JL)   on the E interior
//   JD' = k1D JD + k2D/dt (e~ - e)              Debye,  gD = (1 + k1D) / 2
//   JL' = k1L JL + k3L JLp + k2L/dt (e~ - e);  JLp' = JL      Lorentz
//   e'  = e~ + source (Ez at the TX cell);  trace = e'_z at the probes
// (fdtd_3d_batched.py:753-884, 949-958 with the ADE of fdtd_2d.py:163-178).
// Its transpose, applied from the last step back, carries dL/dtrace to the
// adjoint fields and sums the sensitivities to ca and cb per cell; the host
// turns those into dL/d eps_r and dL/d sigma.  Three launches per step:
// A is pointwise, B needs q at the +1 neighbours, C needs hbar at the -1
// neighbours -- the mirror image of the forward's H-then-E split.
//
// Mapping as the forward kernels: 64 z-cells x 4 y-rows per workgroup,
// blockIdx.z = x-slab + Mx * batch chunk, BB TX planes per thread.
//
// Between two reversed steps C of step s + 1 and A of step s touch the same
// cells only (C reads hbar at the -1 neighbours, then A is pointwise), so A
// can run with C of the step before folded in (WITH_C): lam then crosses HBM
// once per step instead of three times, and a step takes three launches only
// at the ends of the reversed run (A alone first, C alone last).  Contraction
// is off in A, B and C, so both orders give identical results.
#include "fdtd3d_kernels.h"

namespace ceep {

template <int POLES, int BB, bool WITH_C, bool MAGNETIC>
__global__ __launch_bounds__(256) void adj3d_a_kernel(Adj3DArgs p) {
#pragma clang fp contract(off)
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z % p.Mx;
  const int chunk = blockIdx.z / p.Mx;
  if (b >= p.My || d >= p.Mz) return;
  const long long cell = (long long)a * p.sx + (long long)b * p.pitch + d;
  const bool interior = a >= 1 && a <= p.Mx - 2 && b >= 1 && b <= p.My - 2 &&
                        d >= 1 && d <= p.Mz - 2;
  const int b0 = chunk * BB;
  if (!interior) {
    // E never changes off the interior: no adjoint flows through it
#pragma unroll
    for (int k = 0; k < BB; ++k) {
      const int bt = b0 + k;
      if (bt >= p.B) break;
      const long long o = (long long)bt * p.bstride + cell;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        p.lam[c][o] = 0.f;
        p.q[c][o] = 0.f;
      }
    }
    return;
  }
  const float ca = p.ca[cell], cb = p.cb[cell], cbi = cb * p.idx;
  float dk1 = 0.f, dk2 = 0.f, gam = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
  if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; gam = 0.5f * (1.f + dk1); }
  if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
  float gca = 0.f, gcb = 0.f;
  const int nb = min(BB, p.B - b0);
  // every load of the chunk's BB planes first (this kernel is bound by HBM
  // and the loads in flight per wave are what feeds it), then the updates
  float lam[BB][3], aD[BB][3], aL[BB][3], aLp[BB][3], es[BB][3], es1[BB][3];
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    if (k >= nb) break;
    const long long o = (long long)(b0 + k) * p.bstride + cell;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lam[k][c] = p.lam[c][o];
      if (POLES & 1) aD[k][c] = p.ad[c][o];
      if (POLES & 2) { aL[k][c] = p.al[c][o]; aLp[k][c] = p.al_prev[c][o]; }
      es[k][c] = p.e_s[c * p.e_cstride + o];
      es1[k][c] = p.e_s1[c * p.e_cstride + o];
    }
  }
  if (WITH_C) {   // C of step s + 1 (adj3d_c_kernel), then the probe seed of step s
    float c0 = p.ch, cmz = p.ch, cmy = p.ch, cmx = p.ch;
    if (MAGNETIC) {
      c0 = p.dbi[cell];
      cmz = p.dbi[cell - 1];
      cmy = p.dbi[cell - p.pitch];
      cmx = p.dbi[cell - p.sx];
    }
    const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
    const float* __restrict__ hx = p.hbar[0];
    const float* __restrict__ hy = p.hbar[1];
    const float* __restrict__ hz = p.hbar[2];
#pragma unroll
    for (int k = 0; k < BB; ++k) {
      if (k >= nb) break;
      const int bt = b0 + k;
      const long long o = (long long)bt * p.bstride + cell;
      const float rx0 = c0 * hx[o], ry0 = c0 * hy[o], rz0 = c0 * hz[o];
      const float rx_mz = cmz * hx[o - 1], rx_my = cmy * hx[o - p.pitch];
      const float ry_mz = cmz * hy[o - 1], ry_mx = cmx * hy[o - p.sx];
      const float rz_my = cmy * hz[o - p.pitch], rz_mx = cmx * hz[o - p.sx];
      lam[k][0] += (ry0 - ry_mz) - (rz0 - rz_my);
      lam[k][1] += (rz0 - rz_mx) - (rx0 - rx_mz);
      float lz = lam[k][2] + ((rx0 - rx_my) - (ry0 - ry_mx));
      for (int q = probe0; q >= 0; q = p.probe_next[q])
        lz += p.dtr[((long long)bt * p.P + q) * p.T + p.step];
      lam[k][2] = lz;
    }
  }
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    if (k >= nb) break;
    const int bt = b0 + k;
    const long long o = (long long)bt * p.bstride + cell;
    const float inj = cell == p.src_cell[bt] ? p.wave[(long long)bt * p.wave_stride_b + p.step] : 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float G = 0.f;
      if (POLES & 1) G += dk2 * aD[k][c];
      if (POLES & 2) G += lk2 * aL[k][c];
      const float et = lam[k][c] + G;          // adjoint of e~ (it feeds e' and e~ - e)
      p.lam[c][o] = ca * et - G;
      p.q[c][o] = cbi * et;
      if (POLES & 1) p.ad[c][o] = dk1 * aD[k][c] - cb * gam * et;
      if (POLES & 2) p.al_prev[c][o] = lk1 * aL[k][c] + lk3 * aLp[k][c] - cb * et;
      gca += et * es[k][c];
      gcb += et * (es1[k][c] - (c == 2 ? inj : 0.f) - ca * es[k][c]);   // = et cb (curl/dx - sum gJ)
    }
  }
  const long long g = (long long)chunk * p.gstride + cell;
  p.g_ca[g] += gca;
  p.g_cbi[g] += gcb;
}

template <bool MAGNETIC, int BB>
__global__ __launch_bounds__(256) void adj3d_b_kernel(Adj3DArgs p) {
#pragma clang fp contract(off)
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z % p.Mx;
  const int chunk = blockIdx.z / p.Mx;
  if (b >= p.My || d >= p.Mz) return;
  const long long cell = (long long)a * p.sx + (long long)b * p.pitch + d;
  const bool xa = a < p.Mx - 1, yb = b < p.My - 1, zd = d < p.Mz - 1;
  const float da = MAGNETIC ? p.da[cell] : 1.f;
  const float* __restrict__ qx = p.q[0];
  const float* __restrict__ qy = p.q[1];
  const float* __restrict__ qz = p.q[2];
  const int b0 = chunk * BB;
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    const int bt = b0 + k;
    if (bt >= p.B) break;
    const long long o = (long long)bt * p.bstride + cell;
    const float qx0 = qx[o], qy0 = qy[o], qz0 = qz[o];
    // transposes of cy/cz (hx), cx/cz (hy), cx/cy (hz) of the forward E kernel
    if (yb && zd) {
      const float t = (qy0 - qy[o + 1]) - (qz0 - qz[o + p.pitch]);
      p.hbar[0][o] = MAGNETIC ? da * p.hbar[0][o] + t : p.hbar[0][o] + t;
    }
    if (xa && zd) {
      const float t = (qz0 - qz[o + p.sx]) - (qx0 - qx[o + 1]);
      p.hbar[1][o] = MAGNETIC ? da * p.hbar[1][o] + t : p.hbar[1][o] + t;
    }
    if (xa && yb) {
      const float t = (qx0 - qx[o + p.pitch]) - (qy0 - qy[o + p.sx]);
      p.hbar[2][o] = MAGNETIC ? da * p.hbar[2][o] + t : p.hbar[2][o] + t;
    }
  }
}

template <bool MAGNETIC, int BB>
__global__ __launch_bounds__(256) void adj3d_c_kernel(Adj3DArgs p) {
#pragma clang fp contract(off)
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z % p.Mx;
  const int chunk = blockIdx.z / p.Mx;
  if (b >= p.My || d >= p.Mz) return;
  if (!(a >= 1 && a <= p.Mx - 2 && b >= 1 && b <= p.My - 2 && d >= 1 && d <= p.Mz - 2)) return;
  const long long cell = (long long)a * p.sx + (long long)b * p.pitch + d;
  // dbi at this cell and at the -1 neighbours (every referenced H cell lies
  // in its update range because the cell is interior)
  float c0 = p.ch, cmz = p.ch, cmy = p.ch, cmx = p.ch;
  if (MAGNETIC) {
    c0 = p.dbi[cell];
    cmz = p.dbi[cell - 1];
    cmy = p.dbi[cell - p.pitch];
    cmx = p.dbi[cell - p.sx];
  }
  const float* __restrict__ hx = p.hbar[0];
  const float* __restrict__ hy = p.hbar[1];
  const float* __restrict__ hz = p.hbar[2];
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
  const int b0 = chunk * BB;
#pragma unroll
  for (int k = 0; k < BB; ++k) {
    const int bt = b0 + k;
    if (bt >= p.B) break;
    const long long o = (long long)bt * p.bstride + cell;
    const float rx0 = c0 * hx[o], ry0 = c0 * hy[o], rz0 = c0 * hz[o];
    const float rx_mz = cmz * hx[o - 1], rx_my = cmy * hx[o - p.pitch];
    const float ry_mz = cmz * hy[o - 1], ry_mx = cmx * hy[o - p.sx];
    const float rz_my = cmy * hz[o - p.pitch], rz_mx = cmx * hz[o - p.sx];
    p.lam[0][o] += (ry0 - ry_mz) - (rz0 - rz_my);
    p.lam[1][o] += (rz0 - rz_mx) - (rx0 - rx_mz);
    float lz = p.lam[2][o] + ((rx0 - rx_my) - (ry0 - ry_mx));
    if (p.step >= 1)
      for (int q = probe0; q >= 0; q = p.probe_next[q])
        lz += p.dtr[((long long)bt * p.P + q) * p.T + p.step - 1];
    p.lam[2][o] = lz;
  }
}

static dim3 adj_grid3(const Adj3DArgs& a, int bb) {
  return dim3((a.Mz + 63) / 64, (a.My + 3) / 4, a.Mx * ((a.B + bb - 1) / bb));
}

// one reversed step: A (with C of the step before folded in when with_c), then B
template <int POLES, int BB>
static hipError_t launch_adj3d_pb(const Adj3DArgs& a, bool with_c, hipStream_t s) {
  const dim3 blk(64, 4, 1), grd = adj_grid3(a, BB);
  if (a.da) {
    if (with_c) hipLaunchKernelGGL((adj3d_a_kernel<POLES, BB, true, true>), grd, blk, 0, s, a);
    else hipLaunchKernelGGL((adj3d_a_kernel<POLES, BB, false, true>), grd, blk, 0, s, a);
    hipLaunchKernelGGL((adj3d_b_kernel<true, BB>), grd, blk, 0, s, a);
  } else {
    if (with_c) hipLaunchKernelGGL((adj3d_a_kernel<POLES, BB, true, false>), grd, blk, 0, s, a);
    else hipLaunchKernelGGL((adj3d_a_kernel<POLES, BB, false, false>), grd, blk, 0, s, a);
    hipLaunchKernelGGL((adj3d_b_kernel<false, BB>), grd, blk, 0, s, a);
  }
  return hipGetLastError();
}

template <int POLES>
static hipError_t launch_adj3d_p(const Adj3DArgs& a, int bb, bool with_c, hipStream_t s) {
  switch (bb) {
    case 1: return launch_adj3d_pb<POLES, 1>(a, with_c, s);
    case 2: return launch_adj3d_pb<POLES, 2>(a, with_c, s);
    case 4: return launch_adj3d_pb<POLES, 4>(a, with_c, s);
    case 8: return launch_adj3d_pb<POLES, 8>(a, with_c, s);
  }
  return hipErrorInvalidValue;
}

template <int BB>
static hipError_t launch_adj3d_c_b(const Adj3DArgs& a, hipStream_t s) {
  const dim3 blk(64, 4, 1), grd = adj_grid3(a, BB);
  if (a.da) hipLaunchKernelGGL((adj3d_c_kernel<true, BB>), grd, blk, 0, s, a);
  else hipLaunchKernelGGL((adj3d_c_kernel<false, BB>), grd, blk, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adj3d_c(const Adj3DArgs& a, int bb, hipStream_t s) {
  switch (bb) {
    case 1: return launch_adj3d_c_b<1>(a, s);
    case 2: return launch_adj3d_c_b<2>(a, s);
    case 4: return launch_adj3d_c_b<4>(a, s);
    case 8: return launch_adj3d_c_b<8>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_adj3d(const Adj3DArgs& a, int poles, int bb, bool with_c, hipStream_t s) {
  switch (poles) {
    case 0: return launch_adj3d_p<0>(a, bb, with_c, s);
    case 1: return launch_adj3d_p<1>(a, bb, with_c, s);
    case 2: return launch_adj3d_p<2>(a, bb, with_c, s);
    case 3: return launch_adj3d_p<3>(a, bb, with_c, s);
  }
  return hipErrorInvalidValue;
}

__global__ void adj3d_seed_kernel(float* lam_z, const long long* probe_cell, const float* dtr,
                                  int B, int P, int T, int step, long long bstride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * P) return;
  const int bt = i / P, q = i % P;
  const long long pc = probe_cell[q];
  if (pc < 0) return;
  // several probes may share a cell: accumulate atomically
  atomicAdd(lam_z + (long long)bt * bstride + pc, dtr[((long long)bt * P + q) * T + step]);
}

hipError_t launch_adj3d_seed(float* lam_z, const long long* probe_cell, const float* dtr,
                             int B, int P, int T, int step, long long bstride, hipStream_t s) {
  const int n = B * P;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(adj3d_seed_kernel, dim3((n + 255) / 256), dim3(256), 0, s, lam_z,
                     probe_cell, dtr, B, P, T, step, bstride);
  return hipGetLastError();
}

}  // namespace ceep
