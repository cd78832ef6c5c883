// Per-cell coefficient maps and the gradient chain rule of the 2D
// non-dispersive path, formed on the device (gfx950).
//
// The maps are a function of the material maps only, one value per cell, so
// they are formed where the material maps are: in FP64 from the FP64 eps_r /
// sigma maps, rounded once to FP32 (the stencil kernels step FP32 fields).
// With contraction off and IEEE division these are bit for bit the values the
// host forms with the same expressions (fdtd_2d_batched.py:152-157):
//   s = sigma dt / (2 eps_r eps0),  ca = (1 - s) / (1 + s),  cb = (dt / eps) / (1 + s)
// on the E interior of the box (ca = cb = 0 elsewhere, so the stencils need
// no masks).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "coef2d.h"

namespace ceep {
namespace {

constexpr double kEps0d = 8.854187817e-12;   // core/constants.py:24

struct Coef {
  double ca, cb, s, inv_eps;
};

__device__ __forceinline__ Coef coef_of(double eps_r, double sigma, double dt) {
#pragma clang fp contract(off)
  Coef k;
  const double eps = eps_r * kEps0d;
  k.s = sigma * dt / (2.0 * eps);
  k.ca = (1.0 - k.s) / (1.0 + k.s);
  k.cb = (dt / eps) / (1.0 + k.s);
  k.inv_eps = 1.0 / eps;
  return k;
}

}  // namespace

// One thread per box cell (pitch-padded row, padding set to 0).  With a
// reference medium (eps_ref != null) also its maps and the difference-form
// maps d_a = ca - cb ca_ref / cb_ref, d_b = cb / cb_ref - 1 (exactly 0 where
// the two media agree, and off the interior).
__global__ __launch_bounds__(256) void coef2d_kernel(Coef2DArgs p) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int a = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (a >= p.Mx || c >= p.pitch) return;
  const long long o = (long long)a * p.pitch + c;
  const bool inbox = c < p.My;
  const bool interior = p.cpml_interior ? (a >= 1 && c >= 1 && inbox)
                                        : (a >= 1 && a <= p.Mx - 2 && c >= 1 && c <= p.My - 2);
  float ca = 0.f, cb = 0.f, ra = 0.f, rb = 0.f, da = 0.f, db = 0.f;
  if (interior) {
    const long long g = (long long)(a + p.org) * p.ny_full + (c + p.org);
    const Coef k = coef_of(p.eps_r[g], p.sigma[g], p.dt);
    ca = (float)k.ca;
    cb = (float)k.cb;
    if (p.eps_ref) {
      const Coef r = coef_of(p.eps_ref[g], p.sigma_ref[g], p.dt);
      ra = (float)r.ca;
      rb = (float)r.cb;
      const double q = k.cb / r.cb;
      da = (float)(k.ca - q * r.ca);
      db = (float)(q - 1.0);
    }
  }
  p.ca[o] = ca;
  p.cb[o] = cb;
  if (p.eps_ref) {
    p.ca_ref[o] = ra;
    p.cb_ref[o] = rb;
    p.d_a[o] = da;
    p.d_b[o] = db;
  }
}

hipError_t launch_coef2d(const Coef2DArgs& a, hipStream_t s) {
  if (a.Mx < 1 || a.pitch < 1) return hipSuccess;
  const dim3 blk(256, 1, 1), grd((unsigned)((a.pitch + 63) / 64), (unsigned)((a.Mx + 3) / 4), 1);
  hipLaunchKernelGGL(coef2d_kernel, grd, blk, 0, s, a);
  return hipGetLastError();
}

// One workgroup per output tile of the blocked forward launch shape (R rows,
// K steps; tile t covers window rows [tx (R - 2K) - K, + R) and columns
// [ty (64 - 2K) - K, + 64)): flag[t] = 1 if d_a or d_b is nonzero anywhere in
// the window, i.e. if the difference form's source term reaches it.
__global__ __launch_bounds__(256) void diff_tiles_kernel(const float* __restrict__ d_a,
                                                         const float* __restrict__ d_b,
                                                         int Mx, int My, int pitch, int R, int K,
                                                         int tiles_y, unsigned char* flag) {
  const int t = blockIdx.x;
  const int tx = t / tiles_y, ty = t % tiles_y;
  const int a0 = tx * (R - 2 * K) - K, c0 = ty * (64 - 2 * K) - K;
  const int c = c0 + (int)(threadIdx.x & 63);
  int any = 0;
  if (c >= 0 && c < My) {
    for (int r = (int)(threadIdx.x >> 6); r < R; r += 4) {
      const int a = a0 + r;
      if (a < 0 || a >= Mx) continue;
      const long long o = (long long)a * pitch + c;
      any |= (d_a[o] != 0.f) | (d_b[o] != 0.f);
    }
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) flag[t] = any ? 1 : 0;
}

hipError_t launch_diff_tiles(const float* d_a, const float* d_b, int Mx, int My, int pitch,
                             int R, int K, int tiles_x, int tiles_y, unsigned char* flag,
                             hipStream_t s) {
  if (tiles_x < 1 || tiles_y < 1) return hipSuccess;
  hipLaunchKernelGGL(diff_tiles_kernel, dim3((unsigned)(tiles_x * tiles_y)), dim3(256), 0, s,
                     d_a, d_b, Mx, My, pitch, R, K, tiles_y, flag);
  return hipGetLastError();
}

// Chain rule from the adjoint's sums to dL/d eps_r and dL/d sigma on the full
// grid (zero outside the box interior), in FP64:
//   dca/ds = -2 / (1 + s)^2,  ds/deps = -s / eps,  ds/dsigma = dt / (2 eps),
//   dcb/deps / cb = -1/eps - ds/deps / (1 + s),  dcb/dsigma / cb = -ds/dsigma / (1 + s)
// g_cbi holds cb dL/dcb (the adjoint sums lam cb / dx curl), g_ca dL/dca, both
// scaled by 1 / inv_scale.
__global__ __launch_bounds__(256) void chain2d_kernel(Chain2DArgs p) {
#pragma clang fp contract(off)
  const int y = blockIdx.x * 64 + (threadIdx.x & 63);
  const int x = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= p.nx || y >= p.ny) return;
  const long long g = (long long)x * p.ny + y;
  const int a = x - p.org, c = y - p.org;
  double ge = 0.0, gs = 0.0;
  if (a >= 1 && a <= p.Mx - 2 && c >= 1 && c <= p.My - 2) {
    const long long o = (long long)a * p.pitch + c;
    const double dt = p.dt;
    const double eps = p.eps_r[g] * kEps0d, sg = p.sigma[g];
    const double inv_eps = 1.0 / eps;
    const double sv = 0.5 * sg * dt * inv_eps;
    const double inv_1s = 1.0 / (1.0 + sv);
    const double dL_dca = (double)p.g_ca[o] * p.inv_scale;
    const double dL_dcb_cb = (double)p.g_cbi[o] * p.inv_scale;
    const double dca_ds = -2.0 * inv_1s * inv_1s;
    const double ds_deps = -sv * inv_eps, ds_dsig = 0.5 * dt * inv_eps;
    const double rdcb_deps = -inv_eps - inv_1s * ds_deps;
    const double rdcb_dsig = -inv_1s * ds_dsig;
    ge = kEps0d * (dL_dca * dca_ds * ds_deps + dL_dcb_cb * rdcb_deps);
    gs = dL_dca * dca_ds * ds_dsig + dL_dcb_cb * rdcb_dsig;
  }
  p.grad_eps_r[g] = ge;
  p.grad_sigma[g] = gs;
}

hipError_t launch_chain2d(const Chain2DArgs& a, hipStream_t s) {
  if (a.nx < 1 || a.ny < 1) return hipSuccess;
  const dim3 blk(256, 1, 1), grd((unsigned)((a.ny + 63) / 64), (unsigned)((a.nx + 3) / 4), 1);
  hipLaunchKernelGGL(chain2d_kernel, grd, blk, 0, s, a);
  return hipGetLastError();
}

// out[i] = x[i] + y[i] (float4 when n % 4 == 0)
__global__ __launch_bounds__(256) void add4_kernel(const float4* __restrict__ x,
                                                   const float4* __restrict__ y,
                                                   float4* __restrict__ out, long long n4) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x) {
    const float4 u = x[i], v = y[i];
    out[i] = make_float4(u.x + v.x, u.y + v.y, u.z + v.z, u.w + v.w);
  }
}

__global__ __launch_bounds__(256) void add1_kernel(const float* __restrict__ x,
                                                   const float* __restrict__ y,
                                                   float* __restrict__ out, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = x[i] + y[i];
}

hipError_t launch_add(const float* x, const float* y, float* out, long long n, hipStream_t s) {
  if (n < 1) return hipSuccess;
  const int blocks = 2048;
  if (n % 4 == 0 && ((uintptr_t)x | (uintptr_t)y | (uintptr_t)out) % 16 == 0)
    hipLaunchKernelGGL(add4_kernel, dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(x), reinterpret_cast<const float4*>(y),
                       reinterpret_cast<float4*>(out), n / 4);
  else
    hipLaunchKernelGGL(add1_kernel, dim3(blocks), dim3(256), 0, s, x, y, out, n);
  return hipGetLastError();
}

}  // namespace ceep
