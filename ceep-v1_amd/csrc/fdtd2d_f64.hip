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

namespace ceep {
namespace {

// lane l receives lane l+1 (0 for lane 63) / lane l-1 (0 for lane 0): two
// 32-bit DPP wave shifts per double
__device__ __forceinline__ double next_lane_d(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)x, 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), 0x130, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double prev_lane_d(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)x, 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), 0x138, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double bload_d(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bstore_d(double v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  using b64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(b64, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_d(const double* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), 0, (int)bytes, 0x00020000);
}

}  // namespace

// K steps of step2d_f64_kernel on four-wavefront windows.  Cells outside the
// box load as 0 and every E-side coefficient is 0 off the E interior, so the
// uniform update reproduces the one-step kernel's boundary cases exactly:
// hx - ch * (0 - 0) = hx at the last column, E = 0 * e + 0 * (...) = 0 on
// the walls.  Contraction is off and the operand order is that kernel's.
template <int RW, int K>
__global__ __launch_bounds__(256) void step2d_f64_tb_kernel(TB2D64Args p) {
#pragma clang fp contract(off)
  constexpr int R = 4 * RW;
  __shared__ double x_e[4][64], x_h[4][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int task = xcd_block(p.xcd_order);
  if (task >= p.ntasks) return;                       // uniform per workgroup
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = task % tiles;
  const int chunk = task / tiles;
  const int ty = tile % p.tiles_y;
  const int tx = tile / p.tiles_y;
  const int a0 = tx * (R - 2 * K) - K;
  const int cb0 = ty * (64 - 2 * K) - K;
  const int c = cb0 + lane;
  const bool col_ok = c >= 0 && c < p.My;
  const int r0 = w * RW;
  const unsigned pitch8 = (unsigned)p.pitch * 8u;
  const unsigned lane_off = col_ok ? (unsigned)((a0 + r0) * p.pitch + c) * 8u : kOffNone;
  const unsigned plane_bytes = (unsigned)p.Mx * pitch8;
  auto row_off = [&](int i) -> unsigned {
    const int a = a0 + r0 + i;
    return (a >= 0 && a < p.Mx) ? lane_off + (unsigned)i * pitch8 : kOffNone;
  };
  auto is_out = [&](int i) -> bool { return r0 + i >= K && r0 + i < R - K; };
  const bool has_up = w > 0, has_down = w < 3;
  const bool out_lane = lane >= K && lane < 64 - K && col_ok;

  double mca[RW], mcb[RW];
  {
    const __amdgpu_buffer_rsrc_t rca = rsrc_d(p.ca, plane_bytes);
    const __amdgpu_buffer_rsrc_t rcb = rsrc_d(p.cb, plane_bytes);
#pragma unroll
    for (int i = 0; i < RW; ++i) {
      mca[i] = bload_d(rca, row_off(i));
      mcb[i] = bload_d(rcb, row_off(i));
    }
  }
  const double ch = p.ch, idx = p.idx;
  const bool few_probes = p.P <= 64;
  int pr_h = -1, pl_h = -1;
  unsigned long long pmask = 0;
  if (few_probes) {
    const int pc = lane < p.P ? p.probe_cell[lane] : -1;
    int pr_w = -1;
    if (pc >= 0) {
      pr_w = pc / p.pitch - a0;
      pr_h = pr_w - r0;
      pl_h = pc % p.pitch - cb0;
    }
    pmask = __ballot(pc >= 0 && pr_w >= K && pr_w < R - K && pr_h >= 0 && pr_h < RW &&
                     pl_h >= K && pl_h < 64 - K);
  }

  const int b_begin = chunk * p.bpw;
  const int b_end = min(b_begin + p.bpw, p.B);
  for (int bt = b_begin; bt < b_end; ++bt) {
    const long long base = (long long)bt * p.bstride;
    double ez[RW], hx[RW], hy[RW];
    {
      const __amdgpu_buffer_rsrc_t r_ez = rsrc_d(p.ez_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hx = rsrc_d(p.hx_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hy = rsrc_d(p.hy_in + base, plane_bytes);
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const unsigned o = row_off(i);
        ez[i] = bload_d(r_ez, o);
        hx[i] = bload_d(r_hx, o);
        hy[i] = bload_d(r_hy, o);
      }
    }
    const int sc = p.src_cell[bt];
    const int s_row = sc / p.pitch - (a0 + r0);
    const int s_lane = sc % p.pitch - cb0;
    const bool src_here = s_row >= 0 && s_row < RW && s_lane >= 0 && s_lane < 64;

#pragma unroll
    for (int s = 0; s < K; ++s) {
      if (s >= p.ksteps) break;                        // uniform per workgroup
      const int step = p.step0 + s;
      // hx = hx - ch (ez[c+1] - ez[c]),  hy = hy + ch (ez[a+1] - ez[a])
      x_e[w][lane] = ez[0];
      __syncthreads();
      const double e_dn = has_down ? x_e[w + 1][lane] : 0.0;
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const double eR = next_lane_d(ez[i]);
        const double eD = (i + 1 < RW) ? ez[i + 1] : e_dn;
        hx[i] = hx[i] - ch * (eR - ez[i]);
        hy[i] = hy[i] + ch * (eD - ez[i]);
      }
      // ez = ca ez + cb ((hy - hy[a-1]) idx - (hx - hx[c-1]) idx), then the source
      x_h[w][lane] = hy[RW - 1];
      __syncthreads();
      const double h_up = has_up ? x_h[w - 1][lane] : 0.0;
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const double hxl = prev_lane_d(hx[i]);
        const double hyu = (i > 0) ? hy[i - 1] : h_up;
        ez[i] = mca[i] * ez[i] + mcb[i] * ((hy[i] - hyu) * idx - (hx[i] - hxl) * idx);
      }
      if (src_here) {
        const double wv = p.wave[(long long)bt * p.wave_stride_b + step];
#pragma unroll
        for (int i = 0; i < RW; ++i)
          if (i == s_row && lane == s_lane) ez[i] += wv;
      }
      if (p.e_save) {
        float* dst = p.e_save + (long long)s * p.slot + base;
#pragma unroll
        for (int i = 0; i < RW; ++i) {
          const int a = a0 + r0 + i;
          if (out_lane && is_out(i) && a < p.Mx) dst[(long long)a * p.pitch + c] = (float)ez[i];
        }
      }
      for (unsigned long long m = pmask; m;) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int qr = __shfl(pr_h, l);
        const int ql = __shfl(pl_h, l);
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < RW; ++i) v = (i == qr) ? ez[i] : v;
        if (lane == ql) {
          const long long t = ((long long)bt * p.P + l) * p.T + step;
          p.traces[t] = v;
          p.traces32[t] = (float)v;
        }
      }
      for (int q0 = 0; !few_probes && q0 < p.P; q0 += 64) {
        const int q = q0 + lane;
        const int pc = q < p.P ? p.probe_cell[q] : -1;
        int pw = -1, pr = -1, pl = -1;
        if (pc >= 0) {
          pw = pc / p.pitch - a0;
          pr = pw - r0;
          pl = pc % p.pitch - cb0;
        }
        const bool mine = pc >= 0 && pw >= K && pw < R - K && pr >= 0 && pr < RW &&
                          pl >= K && pl < 64 - K;
        unsigned long long m = __ballot(mine);
        while (m) {
          const int l = __ffsll((long long)m) - 1;
          m &= m - 1;
          const int qr = __shfl(pr, l);
          const int ql = __shfl(pl, l);
          double v = 0.0;
#pragma unroll
          for (int i = 0; i < RW; ++i) v = (i == qr) ? ez[i] : v;
          if (lane == ql) {
            const long long t = ((long long)bt * p.P + q0 + l) * p.T + step;
            p.traces[t] = v;
            p.traces32[t] = (float)v;
          }
        }
      }
    }
    {
      const __amdgpu_buffer_rsrc_t r_ez = rsrc_d(p.ez_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hx = rsrc_d(p.hx_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hy = rsrc_d(p.hy_out + base, plane_bytes);
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const unsigned o = (out_lane && is_out(i)) ? row_off(i) : kOffNone;
        bstore_d(ez[i], r_ez, o);
        bstore_d(hx[i], r_hx, o);
        bstore_d(hy[i], r_hy, o);
      }
    }
  }
}

template <int RW>
static hipError_t launch_f64_tb_r(const TB2D64Args& a, int K, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((unsigned)a.ntasks, 1, 1);
  switch (K) {
    case 2: hipLaunchKernelGGL((step2d_f64_tb_kernel<RW, 2>), grd, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((step2d_f64_tb_kernel<RW, 4>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_step2d_f64_tb(const TB2D64Args& a, int R, int K, hipStream_t s) {
  if (a.ntasks < 1) return hipSuccess;
  switch (R) {
    case 16: return launch_f64_tb_r<4>(a, K, s);
    case 32: return launch_f64_tb_r<8>(a, K, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ceep
