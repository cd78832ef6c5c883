// Trace -> S-matrix DFT and field-energy reduction (gfx950).
//
// DFT: S[f][rx][tx] = dt * sum_s trace[tx][rx][s] * exp(-i w_f s dt), the sum of
// DFTMonitor.update over the run (src/ceep/solvers/dft.py:75-90 with the
// pre-increment step index of fdtd_2d.py:287-289), laid out as
// MultistaticSParameters.s_matrix (s_params.py:70-99).  One thread owns one
// (f, rx, tx) output and keeps its two Kahan-compensated FP32 accumulators in
// registers; consecutive lanes take consecutive frequencies, so the twiddle
// table [T][F] is read coalesced and the trace sample is a wave-uniform load.
#include <hip/hip_runtime.h>
#include "reduce_kernels.h"

#include <algorithm>

namespace ceep {

__global__ __launch_bounds__(256) void dft_kernel(const float* __restrict__ traces,
                                                  const float2* __restrict__ tw, int B,
                                                  int P, int T, int F, float dt,
                                                  float2* __restrict__ S) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int pair = blockIdx.y;            // tx * P + rx
  if (f >= F) return;
  const int tx = pair / P, rx = pair % P;
  const float* __restrict__ tr = traces + (long long)pair * T;
  float sr = 0.f, si = 0.f, cr = 0.f, ci = 0.f;
  // unrolled so the trace and twiddle loads of several samples are in flight
  // together (the sums keep their sample order)
#pragma unroll 32
  for (int s = 0; s < T; ++s) {
    const float v = tr[s];
    const float2 w = tw[(long long)s * F + f];
    // Kahan: the product is formed first, then compensated accumulation
    const float yr = __fmul_rn(w.x, v) - cr;
    const float tr_ = sr + yr;
    cr = (tr_ - sr) - yr;
    sr = tr_;
    const float yi = __fmul_rn(w.y, v) - ci;
    const float ti = si + yi;
    ci = (ti - si) - yi;
    si = ti;
  }
  S[((long long)f * P + rx) * B + tx] = make_float2(sr * dt, si * dt);
}

hipError_t launch_dft(const float* traces, const float2* tw, int B, int P, int T, int F,
                      float dt, float2* S, hipStream_t s) {
  dim3 blk(64, 1, 1);
  dim3 grd((F + 63) / 64, B * P, 1);
  hipLaunchKernelGGL(dft_kernel, grd, blk, 0, s, traces, tw, B, P, T, F, dt, S);
  return hipGetLastError();
}

// Sum of squares of `n` floats; one partial per workgroup (host sums them in FP64).
__global__ __launch_bounds__(256) void sumsq_kernel(const float* __restrict__ x, long long n,
                                                    float* __restrict__ partial) {
  __shared__ float red[256];
  float acc = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Device-to-device copy of n floats (both 16-byte aligned, as every buffer of
// the engine is); a kernel keeps the copy on the shader engines at HBM speed.
__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, long long n4) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy(const float* src, float* dst, long long n, hipStream_t s) {
  if (n % 4 != 0) return hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s);
  const long long n4 = n / 4;
  const int blocks = (int)std::min<long long>((n4 + 255) / 256, 8192);
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void zero_kernel(unsigned* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
  const long long n = (long long)(bytes / 4);
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, static_cast<unsigned*>(p), n);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void to_f32_kernel(const double* __restrict__ src,
                                                     float* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

hipError_t launch_to_f32(const double* src, float* dst, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(to_f32_kernel, dim3(blocks), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_sumsq(const float* x, long long n, float* partial, int nblocks,
                        hipStream_t s) {
  hipLaunchKernelGGL(sumsq_kernel, dim3(nblocks), dim3(256), 0, s, x, n, partial);
  return hipGetLastError();
}

// Twiddle table [T][F] = (cos(w_f s dt), -sin(w_f s dt)), w_f = 2 pi f, formed
// in FP64 and rounded once (the host's table for ceep_fdtd_dft uses the same
// expressions).
__global__ __launch_bounds__(256) void twiddle_kernel(const double* __restrict__ freqs, int F,
                                                      int T, double dt, float2* __restrict__ tw) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)T * F) return;
  const int s = (int)(i / F), f = (int)(i % F);
  const double t = s * dt;
  const double w = 2.0 * 3.14159265358979323846 * freqs[f];
  tw[i] = make_float2((float)cos(w * t), (float)(-sin(w * t)));
}

hipError_t launch_twiddles(const double* freqs, int F, int T, double dt, float2* tw,
                           hipStream_t s) {
  const long long n = (long long)T * F;
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, freqs, F,
                     T, dt, tw);
  return hipGetLastError();
}

// out[2i], out[2i+1] = (double)S[i].x, (double)S[i].y
__global__ __launch_bounds__(256) void c2z_kernel(const float2* __restrict__ S, double* __restrict__ out,
                                                  long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = S[i];
  out[2 * i] = (double)v.x;
  out[2 * i + 1] = (double)v.y;
}

hipError_t launch_c2z(const float2* S, double* out, long long n, hipStream_t s) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(c2z_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, out, n);
  return hipGetLastError();
}

}  // namespace ceep
