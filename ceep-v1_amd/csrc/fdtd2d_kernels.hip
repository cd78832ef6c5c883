// Batched 2D TMz step kernels for gfx950.
//
// One launch advances every TX of the batch by one full leapfrog step:
//   H half step  (fdtd_2d_batched.py:472-483; run_cpu form :574-583)
//   E full step  (:492-498) with the ADE of fdtd_2d.py:163-178 /
//                dispersive.py:341-404 when poles are present
//   source injection (:509-515) and probe gathering (:517-526)
// fused, reading state buffer `in` and writing buffer `out` (ping-pong), so
// the neighbour H values the E update needs are recomputed from `in`
// instead of being re-read after a device-wide barrier.
//
// Mapping: a 64x4 workgroup covers 64 consecutive y-cells (one wavefront
// per x-row, fully coalesced 256 B row segments) and 4 x-rows; each thread
// loops over a chunk of BB batch planes so the per-cell material
// coefficients are read once per chunk instead of once per TX.
#include "fdtd_common.h"
#include "fdtd2d_kernels.h"

namespace ceep {

template <int POLES, int BB>
__global__ __launch_bounds__(256) void step2d_kernel(Step2DArgs p) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  const int a = blockIdx.y * 4 + threadIdx.y;
  if (a >= p.Mx || c >= p.My) return;
  const long long cell = (long long)a * p.pitch + c;
  const bool interior = (a >= 1) && (a <= p.Mx - 2) && (c >= 1) && (c <= p.My - 2);
  const bool has_r = c < p.My - 1;   // hx(a,c) is updated (reads ez(a,c+1))
  const bool has_d = a < p.Mx - 1;   // hy(a,c) is updated (reads ez(a+1,c))

  float ca = 0.f, cb = 0.f;
  float dk1 = 0.f, dk2 = 0.f, lk1 = 0.f, lk2 = 0.f, lk3 = 0.f;
  if (interior) {
    ca = p.ca[cell];
    cb = p.cb[cell];
    if (POLES & 1) { dk1 = p.d_k1[cell]; dk2 = p.d_k2dt[cell]; }
    if (POLES & 2) { lk1 = p.l_k1[cell]; lk2 = p.l_k2dt[cell]; lk3 = p.l_k3[cell]; }
  }
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
  const float ch = p.ch;
  const float idx = p.idx;

  const int b0 = blockIdx.z * BB;
#pragma unroll
  for (int bb = 0; bb < BB; ++bb) {
    const int bt = b0 + bb;
    if (bt >= p.B) break;
    const long long o = (long long)bt * p.bstride + cell;
    const float* __restrict__ ez = p.ez_in + o;
    const float* __restrict__ hx = p.hx_in + o;
    const float* __restrict__ hy = p.hy_in + o;

    const float e0 = ez[0];
    const float eR = has_r ? ez[1] : 0.f;
    const float eD = has_d ? ez[p.pitch] : 0.f;
    // H half step at this cell (run_cpu form: cells on the last row/column keep 0)
    const float hxn = has_r ? h_minus(hx[0], ch, eR - e0) : hx[0];
    const float hyn = has_d ? h_plus(hy[0], ch, eD - e0) : hy[0];
    p.hx_out[o] = hxn;
    p.hy_out[o] = hyn;

    float en = 0.f;     // box walls: zero before injection (fdtd_2d_batched.py:504-507)
    if (interior) {
      // neighbour H values at (a, c-1) and (a-1, c), recomputed bit-identically
      const float eL = ez[-1];
      const float eU = ez[-p.pitch];
      const float hxl = h_minus(hx[-1], ch, e0 - eL);
      const float hyu = h_plus(hy[-p.pitch], ch, e0 - eU);
      const float curl = (hyn - hyu) - (hxn - hxl);
      if (POLES == 0) {
        en = e_step(ca, cb, e0, curl, idx);
      } else {
        float jd = 0.f, jl = 0.f, jlp = 0.f, sgj = 0.f;
        if (POLES & 1) {
          jd = p.jd[o];
          sgj += 0.5f * (1.f + dk1) * jd;          // gamma = (1 + k1) / 2
        }
        if (POLES & 2) {
          jl = p.jl_cur[o];
          jlp = p.jl_prev[o];
          sgj += jl;                                // gamma = 1
        }
        en = __builtin_fmaf(-cb, sgj, e_step(ca, cb, e0, curl, idx));
        const float de = en - e0;
        if (POLES & 1) p.jd[o] = j_debye(dk1, dk2, jd, de);
        if (POLES & 2) p.jl_prev[o] = lk1 * jl + lk3 * jlp + lk2 * de;  // becomes J_cur
      }
    }
    if (cell == p.src_cell[bt]) en += p.wave[(long long)bt * p.wave_stride_b + p.step];
    p.ez_out[o] = en;
    if (p.e_save) p.e_save[o] = en;
    for (int q = probe0; q >= 0; q = p.probe_next[q])
      p.traces[((long long)bt * p.P + q) * p.T + p.step] = en;
  }
}

template <int POLES>
static hipError_t launch_pole(const Step2DArgs& a, int bb, hipStream_t s) {
  dim3 blk(64, 4, 1);
  dim3 grd((a.My + 63) / 64, (a.Mx + 3) / 4, (a.B + bb - 1) / bb);
  switch (bb) {
    case 1: hipLaunchKernelGGL((step2d_kernel<POLES, 1>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((step2d_kernel<POLES, 2>), grd, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((step2d_kernel<POLES, 4>), grd, blk, 0, s, a); break;
    case 8: hipLaunchKernelGGL((step2d_kernel<POLES, 8>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_step2d(const Step2DArgs& a, int poles, int bb, hipStream_t s) {
  switch (poles) {
    case 0: return launch_pole<0>(a, bb, s);
    case 1: return launch_pole<1>(a, bb, s);
    case 2: return launch_pole<2>(a, bb, s);
    case 3: return launch_pole<3>(a, bb, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ceep
