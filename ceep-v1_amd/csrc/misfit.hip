// S-parameter misfit L = sum over (f, rx, tx) of |S - S_target|^2 and its
// gradient dL/dtrace on the device, for the 2D gradient path (gfx950).
//
// S[f][rx][tx] = dt sum_s trace[tx][rx][s] exp(-i w_f s dt) is the running DFT
// of src/ceep/solvers/dft.py:75-90 in the S layout of s_params.py:70-99; the
// loss is the misfit the gradient path differentiates (the reference has no
// adjoint: src/ceep/ai/__init__.py).  Sums run in FP64 from the FP32 traces,
// as the host path does after widening the traces to FP64.
#include <hip/hip_runtime.h>
#include "misfit.h"

namespace ceep {

// One thread per (f, rx, tx): lanes take consecutive frequencies (the twiddle
// row is read coalesced, the trace sample is wave-uniform).
__global__ __launch_bounds__(64) void misfit_s_kernel(const float* __restrict__ traces,
                                                      const double2* __restrict__ tw64, int B,
                                                      int P, int T, int F, double dt,
                                                      const double2* __restrict__ s_tgt,
                                                      double2* __restrict__ D) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int pair = blockIdx.y;            // tx * P + rx
  if (f >= F) return;
  const int tx = pair / P, rx = pair % P;
  const float* __restrict__ tr = traces + (long long)pair * T;
  double sr = 0.0, si = 0.0;
#pragma unroll 16
  for (int s = 0; s < T; ++s) {
    const double v = (double)tr[s];
    const double2 w = tw64[(long long)s * F + f];
    sr += w.x * v;
    si += w.y * v;
  }
  const long long o = ((long long)f * P + rx) * B + tx;
  double2 d = make_double2(sr * dt, si * dt);
  if (s_tgt) {
    d.x -= s_tgt[o].x;
    d.y -= s_tgt[o].y;
  }
  D[o] = d;
}

__global__ __launch_bounds__(256) void misfit_sum_kernel(const double2* __restrict__ D, int B,
                                                         int P, int F, double dt,
                                                         double* __restrict__ out) {
  __shared__ double s_sum[256], s_max[256];
  const int PB = P * B;
  double acc = 0.0, mx = 0.0;
  for (int pb = threadIdx.x; pb < PB; pb += blockDim.x) {
    double bound = 0.0;
    for (int f = 0; f < F; ++f) {
      const double2 d = D[(long long)f * PB + pb];
      acc += d.x * d.x + d.y * d.y;
      bound += fabs(d.x) + fabs(d.y);
    }
    mx = fmax(mx, 2.0 * dt * bound);
  }
  s_sum[threadIdx.x] = acc;
  s_max[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_sum[0];
    out[1] = s_max[0];
  }
}

// One thread per (tx, rx, s); a workgroup covers 256 samples of one trace.
__global__ __launch_bounds__(256) void misfit_grad_kernel(const double2* __restrict__ D,
                                                          const double2* __restrict__ tw64,
                                                          int B, int P, int T, int F,
                                                          double c, float* __restrict__ dtr) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int pair = blockIdx.y;            // tx * P + rx
  if (s >= T) return;
  const int tx = pair / P, rx = pair % P;
  const double2* __restrict__ w = tw64 + (long long)s * F;
  double g = 0.0;
  for (int f = 0; f < F; ++f) {
    const double2 d = D[((long long)f * P + rx) * B + tx];
    g += d.x * w[f].x + d.y * w[f].y;
  }
  dtr[(long long)pair * T + s] = (float)(c * g);
}

hipError_t launch_misfit_s(const float* traces, const double2* tw64, int B, int P, int T, int F,
                           double dt, const double2* s_tgt, double2* D, hipStream_t s) {
  const dim3 blk(64, 1, 1), grd((unsigned)((F + 63) / 64), (unsigned)(B * P), 1);
  hipLaunchKernelGGL(misfit_s_kernel, grd, blk, 0, s, traces, tw64, B, P, T, F, dt, s_tgt, D);
  return hipGetLastError();
}

hipError_t launch_misfit_sum(const double2* D, int B, int P, int F, double dt, double* out,
                             hipStream_t s) {
  hipLaunchKernelGGL(misfit_sum_kernel, dim3(1), dim3(256), 0, s, D, B, P, F, dt, out);
  return hipGetLastError();
}

hipError_t launch_misfit_grad(const double2* D, const double2* tw64, int B, int P, int T, int F,
                              double dt, double scale, float* dtr, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((unsigned)((T + 255) / 256), (unsigned)(B * P), 1);
  hipLaunchKernelGGL(misfit_grad_kernel, grd, blk, 0, s, D, tw64, B, P, T, F, 2.0 * dt * scale,
                     dtr);
  return hipGetLastError();
}

}  // namespace ceep
