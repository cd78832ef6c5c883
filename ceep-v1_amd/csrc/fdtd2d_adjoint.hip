// Discrete adjoint of the batched 2D TMz update (non-dispersive media).
//
// The forward step of fdtd2d_kernels.hip, written on the stepped box, is
//   hx' = hx - ch Gy(e),  hy' = hy + ch Gx(e),
//   e'  = ca e + cbi (Dx hy' - Dy hx') + w(s) at the source,  trace(s) = e'(probe)
// with Gx e(a) = e(a+1) - e(a), Dx h(a) = h(a) - h(a-1) (likewise along y) and
// ca = cbi = 0 off the E interior.  Its transpose, run from s = T-1 down to 0,
// is the leapfrog below with the roles of E and H swapped; it needs the
// forward E at steps s and s+1 only for the parameter gradients.  One launch
// per adjoint step, ping-pong buffers, one thread per cell looping over every
// TX so the per-cell gradient sums need no atomics.
#include "fdtd_common.h"
#include "fdtd2d_kernels.h"

namespace ceep {

__global__ __launch_bounds__(256) void adj2d_kernel(Adj2DArgs p) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  const int a = blockIdx.y * 4 + threadIdx.y;
  if (a >= p.Mx || c >= p.My) return;
  const long long cell = (long long)a * p.pitch + c;
  const bool hasR = c + 1 < p.My, hasD = a + 1 < p.Mx, hasL = c >= 1, hasU = a >= 1;
  const float ca0 = p.ca[cell];
  const float b0 = p.cb[cell] * p.idx;
  const float bR = hasR ? p.cb[cell + 1] * p.idx : 0.f;
  const float bD = hasD ? p.cb[cell + p.pitch] * p.idx : 0.f;
  const float bL = hasL ? p.cb[cell - 1] * p.idx : 0.f;
  const float bU = hasU ? p.cb[cell - p.pitch] * p.idx : 0.f;
  const int probe0 = p.probe_head[cell];
  const float ch = p.ch;
  float gca = 0.f, gcb = 0.f;
  for (int bt = 0; bt < p.B; ++bt) {
    const long long o = (long long)bt * p.bstride + cell;
    const float* le = p.le_in + o;
    const float l0 = le[0];
    const float q0 = b0 * l0;
    const float qR = hasR ? bR * le[1] : 0.f;
    const float qD = hasD ? bD * le[p.pitch] : 0.f;
    const float qL = hasL ? bL * le[-1] : 0.f;
    const float qU = hasU ? bU * le[-p.pitch] : 0.f;
    const float mx = p.mx_in[o] + (qR - q0);                       // mu_x'(a, c)
    const float my = p.my_in[o] + (q0 - qD);                       // mu_y'(a, c)
    const float mxL = hasL ? p.mx_in[o - 1] + (q0 - qL) : 0.f;     // mu_x'(a, c-1)
    const float myU = hasU ? p.my_in[o - p.pitch] + (qU - q0) : 0.f;  // mu_y'(a-1, c)
    p.mx_out[o] = mx;
    p.my_out[o] = my;
    float ln = ca0 * l0 + ch * ((myU - my) - (mxL - mx));
    if (p.step >= 1)
      for (int q = probe0; q >= 0; q = p.probe_next[q])
        ln += p.dtr[((long long)bt * p.P + q) * p.T + (p.step - 1)];
    p.le_out[o] = ln;
    if (b0 != 0.f) {
      const float es = p.e_s[o];
      float d = p.e_s1[o] - ca0 * es;
      if (cell == p.src_cell[bt]) d -= p.wave[(long long)bt * p.wave_stride_b + p.step];
      gca += l0 * es;
      gcb += l0 * d;
    }
  }
  if (b0 != 0.f) {
    p.g_ca[cell] += gca;
    p.g_cbi[cell] += gcb;
  }
}

hipError_t launch_adj2d(const Adj2DArgs& a, hipStream_t s) {
  dim3 blk(64, 4, 1), grd((a.My + 63) / 64, (a.Mx + 3) / 4, 1);
  hipLaunchKernelGGL(adj2d_kernel, grd, blk, 0, s, a);
  return hipGetLastError();
}

__global__ void adj_seed_kernel(float* le, const int* probe_cell, const float* dtr, int B, int P,
                                int T, int step, long long bstride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * P) return;
  const int bt = i / P, q = i % P;
  const int pc = probe_cell[q];
  if (pc < 0) return;
  // several probes may share a cell: accumulate atomically
  atomicAdd(le + (long long)bt * bstride + pc, dtr[((long long)bt * P + q) * T + step]);
}

__global__ __launch_bounds__(256) void sum_parts_kernel(const float* __restrict__ src, int nparts,
                                                        long long stride, long long n,
                                                        float* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.f;
  for (int k = 0; k < nparts; ++k) acc += src[k * stride + i];
  dst[i] = acc;
}

hipError_t launch_sum_parts(const float* src, int nparts, long long stride, long long n,
                            float* dst, hipStream_t s) {
  hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src,
                     nparts, stride, n, dst);
  return hipGetLastError();
}

hipError_t launch_adj_seed(float* le, const int* probe_cell, const float* dtr, int B, int P,
                           int T, int step, long long bstride, hipStream_t s) {
  const int n = B * P;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(adj_seed_kernel, dim3((n + 255) / 256), dim3(256), 0, s, le, probe_cell, dtr,
                     B, P, T, step, bstride);
  return hipGetLastError();
}

}  // namespace ceep
