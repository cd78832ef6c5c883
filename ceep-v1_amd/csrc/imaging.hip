// Delay-and-sum (DAS) and delay-multiply-and-sum (DMAS) imaging from a
// multistatic S-matrix on gfx950 (the reference's DelayAndSumBeamformer,
// src/ceep/imaging/beamforming.py:86-313).
//
// One thread per pixel, FP64 throughout (the work is a small compute-bound
// sum, so the image agrees with the reference's FP64 result to rounding):
// for every (TX, RX) pair in the reference's order the two-way delay
// tau = (|p - tx| + |p - rx|) / v and the pair signal
// sum_f S[f][rx][tx] exp(-i 2 pi f tau); DAS sums the signals, DMAS sums
// s_k1 conj(s_k2) over pairs k1 < k2 as sum_k P_k conj(s_k) with the running
// prefix P_k = sum_{k1 < k} s_k1 -- O(pairs) instead of O(pairs^2).  The
// S-matrix is read by every thread in the same order (L1/L2 broadcast).
#include <hip/hip_runtime.h>

#include "imaging.h"

namespace ceep {

__global__ __launch_bounds__(256) void beamform_kernel(BeamformArgs p) {
  const long long pix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= (long long)p.nx * p.ny) return;
  const int i = (int)(pix / p.ny), j = (int)(pix % p.ny);
  const double px = (p.x0 + i) * p.dx, py = (p.y0 + j) * p.dx;
  const int N = p.n_ant;
  double ar = 0.0, ai = 0.0, pr = 0.0, pi = 0.0;
  for (int t = 0; t < N; ++t) {
    const double tx = p.ant[2 * t], ty = p.ant[2 * t + 1];
    const double d_tx = sqrt((px - tx) * (px - tx) + (py - ty) * (py - ty));
    for (int r = 0; r < N; ++r) {
      const double rx = p.ant[2 * r], ry = p.ant[2 * r + 1];
      const double d_rx = sqrt((px - rx) * (px - rx) + (py - ry) * (py - ry));
      const double delay = (d_tx + d_rx) / p.velocity;
      double sr = 0.0, si = 0.0;
      for (int f = 0; f < p.n_freq; ++f) {
        double sn, cs;
        sincos(p.two_pi * p.freqs[f] * delay, &sn, &cs);
        const double2 s = p.s[((long long)f * N + r) * N + t];
        // s * exp(-i phase) = (s.x + i s.y)(cs - i sn)
        sr += s.x * cs + s.y * sn;
        si += s.y * cs - s.x * sn;
      }
      if (p.dmas) {
        // + P * conj(s)
        ar += pr * sr + pi * si;
        ai += pi * sr - pr * si;
        pr += sr;
        pi += si;
      } else {
        ar += sr;
        ai += si;
      }
    }
  }
  p.out[pix] = sqrt(ar * ar + ai * ai);
}

hipError_t launch_beamform(const BeamformArgs& a, hipStream_t s) {
  const long long n = (long long)a.nx * a.ny;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(beamform_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace ceep
