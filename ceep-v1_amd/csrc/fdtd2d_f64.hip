// FP64 one-step 2D TMz kernel for the gradient path's forward pass (gfx950).
//
// Same fused step as step2d_kernel (H half step, E step, walls, injection,
// probes; fdtd_2d_batched.py:472-526 in the run_cpu form :574-603), for
// non-dispersive media, on FP64 fields.  Misfits between S-matrices that
// differ by a small scattered signal (a TX = RX probe sees the source pulse
// itself) need trace accuracy beyond FP32; this kernel provides it for the
// forward and recomputation passes while the adjoint stays FP32.  Floating-
// point contraction is off and every expression keeps the reference's
// operand order, so results are bitwise those of an FP64 CPU evaluation of
// the same update.  E after the step is also written, rounded to FP32, to
// the adjoint's E store and to the FP32 trace buffer.
#include "fdtd2d_kernels.h"

namespace ceep {

template <int BB>
__global__ __launch_bounds__(256) void step2d_f64_kernel(Step2D64Args p) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * 64 + threadIdx.x;
  const int a = blockIdx.y * 4 + threadIdx.y;
  if (a >= p.Mx || c >= p.My) return;
  const long long cell = (long long)a * p.pitch + c;
  const bool interior = (a >= 1) && (a <= p.Mx - 2) && (c >= 1) && (c <= p.My - 2);
  const bool has_r = c < p.My - 1;
  const bool has_d = a < p.Mx - 1;
  const double ca = interior ? p.ca[cell] : 0.0;
  const double cb = interior ? p.cb[cell] : 0.0;
  const int probe0 = p.probe_head ? p.probe_head[cell] : -1;
  const double ch = p.ch, idx = p.idx;
  const int b0 = blockIdx.z * BB;
#pragma unroll
  for (int bb = 0; bb < BB; ++bb) {
    const int bt = b0 + bb;
    if (bt >= p.B) break;
    const long long o = (long long)bt * p.bstride + cell;
    const double* __restrict__ ez = p.ez_in + o;
    const double* __restrict__ hx = p.hx_in + o;
    const double* __restrict__ hy = p.hy_in + o;
    const double e0 = ez[0];
    // hx = da hx - (db / dx)(ez[c+1] - ez[c]) with da = 1; hy likewise (+)
    const double hxn = has_r ? hx[0] - ch * (ez[1] - e0) : hx[0];
    const double hyn = has_d ? hy[0] + ch * (ez[p.pitch] - e0) : hy[0];
    p.hx_out[o] = hxn;
    p.hy_out[o] = hyn;
    double en = 0.0;
    if (interior) {
      const double hxl = hx[-1] - ch * (e0 - ez[-1]);
      const double hyu = hy[-p.pitch] + ch * (e0 - ez[-p.pitch]);
      // ca ez + cb ((hy - hy[a-1]) / dx - (hx - hx[c-1]) / dx)   (:492-498)
      en = ca * e0 + cb * ((hyn - hyu) * idx - (hxn - hxl) * idx);
    }
    if (cell == p.src_cell[bt]) en += p.wave[(long long)bt * p.wave_stride_b + p.step];
    p.ez_out[o] = en;
    if (p.e_save) p.e_save[o] = (float)en;
    for (int q = probe0; q >= 0; q = p.probe_next[q]) {
      const long long t = ((long long)bt * p.P + q) * p.T + p.step;
      p.traces[t] = en;
      p.traces32[t] = (float)en;
    }
  }
}

hipError_t launch_step2d_f64(const Step2D64Args& a, hipStream_t s) {
  constexpr int BB = 4;
  dim3 blk(64, 4, 1);
  dim3 grd((a.My + 63) / 64, (a.Mx + 3) / 4, (a.B + BB - 1) / BB);
  hipLaunchKernelGGL((step2d_f64_kernel<BB>), grd, blk, 0, s, a);
  return hipGetLastError();
}

}  // namespace ceep
