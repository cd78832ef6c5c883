// Temporally blocked batched 2D TMz step kernel for gfx950.
//
// Same update as step2d_kernel (fdtd_2d_batched.py:472-498, 509-526 with the
// ADE of fdtd_2d.py:163-178), but one launch advances the state by up to K
// steps.  A wavefront owns an R x 64 window of one batch plane: lane l holds
// column c = c0 - K + l, register slot r holds row a = a0 - K + r.  The
// y-neighbours come from the adjacent lanes through DPP wave shifts, the
// x-neighbours are the adjacent register slots, so a step needs neither LDS
// nor a barrier.  Every step the window's outer ring becomes stale (its
// neighbours are outside the window), so after K steps the inner
// (R - 2K) x (64 - 2K) cells are exact and are the only ones written back;
// neighbouring wavefronts recompute the overlap.  Traffic per cell-step drops
// from one read + one write of the state to about (R*64/((R-2K)(64-2K)) + 1)/K.
//
// Cells outside the E interior carry ca = cb = k1 = k2 = 0 and cells outside
// the box are loaded as 0, so the uniform update leaves them at zero (before
// a wall source is added, as in run_cpu) without per-cell masks.
#include "fdtd_common.h"
#include "fdtd2d_kernels.h"

namespace ceep {
namespace {

// lane l receives lane l+1 (0 for lane 63): the value at column c+1
__device__ __forceinline__ float from_next_lane(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
// lane l receives lane l-1 (0 for lane 0): the value at column c-1
__device__ __forceinline__ float from_prev_lane(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

}  // namespace

// One task = one (output tile, batch chunk) window advanced p.ksteps steps by
// the calling wavefront.  Only the rows that can still reach the output
// window are updated: at local step k (0-based) H rows [k, R-k-1) and E rows
// [k+1, R-k-1) (the rest of the window is stale and never read again), which
// skips about 30 % of the row updates at R = 16, K = 4.  Loads and stores go
// through buffer resources with one byte offset per window row.
template <int R, int K, int POLES>
__device__ __forceinline__ void tb_task(const TB2DArgs& p, int task, int lane) {
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = task % tiles;
  const int chunk = task / tiles;
  const int ty = tile % p.tiles_y;
  const int tx = tile / p.tiles_y;
  const int a0 = tx * (R - 2 * K) - K;                // row of register slot 0
  const int cb0 = ty * (64 - 2 * K) - K;              // column of lane 0
  const int c = cb0 + lane;
  const bool col_ok = c >= 0 && c < p.My;
  // byte offset of this lane's cell in window row r; kOffNone off the box
  // (a0 may be negative: the unsigned sum wraps back on rows inside the box)
  const unsigned pitch4 = (unsigned)p.pitch * 4u;
  const unsigned lane_off = col_ok ? (unsigned)(a0 * p.pitch + c) * 4u : kOffNone;
  const unsigned plane_bytes = (unsigned)p.Mx * pitch4;
  auto row_off = [&](int r) -> unsigned {
    const int a = a0 + r;
    return (a >= 0 && a < p.Mx) ? lane_off + (unsigned)r * pitch4 : kOffNone;
  };

  float mca[R], mcb[R], mk1[R], mk2[R];
  {
    const __amdgpu_buffer_rsrc_t rca = plane_rsrc(p.ca, plane_bytes);
    const __amdgpu_buffer_rsrc_t rcb = plane_rsrc(p.cb, plane_bytes);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      mca[r] = bload(rca, row_off(r));
      mcb[r] = bload(rcb, row_off(r));
    }
    if (POLES & 1) {
      const __amdgpu_buffer_rsrc_t rk1 = plane_rsrc(p.d_k1, plane_bytes);
      const __amdgpu_buffer_rsrc_t rk2 = plane_rsrc(p.d_k2dt, plane_bytes);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        mk1[r] = bload(rk1, row_off(r));
        mk2[r] = bload(rk2, row_off(r));
      }
    }
  }
  const float ch = p.ch, idx = p.idx;
  const bool out_lane = lane >= K && lane < 64 - K && col_ok;
  // probes on this wavefront's output window (lane q describes probe q when
  // P <= 64; larger probe sets are scanned every step)
  const bool few_probes = p.P <= 64;
  int pr_h = -1, pl_h = -1;
  unsigned long long pmask = 0;
  if (few_probes) {
    const int pc = lane < p.P ? p.probe_cell[lane] : -1;
    if (pc >= 0) {
      pr_h = pc / p.pitch - a0;
      pl_h = pc % p.pitch - cb0;
    }
    pmask = __ballot(pc >= 0 && pr_h >= K && pr_h < R - K && pl_h >= K && pl_h < 64 - K);
  }

  const int b_begin = chunk * p.bpw;
  const int b_end = min(b_begin + p.bpw, p.B);
  for (int bt = b_begin; bt < b_end; ++bt) {
    const long long base = (long long)bt * p.bstride;
    float ez[R], hx[R], hy[R], jd[R];
    {
      const __amdgpu_buffer_rsrc_t r_ez = plane_rsrc(p.ez_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hx = plane_rsrc(p.hx_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hy = plane_rsrc(p.hy_in + base, plane_bytes);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const unsigned o = row_off(r);
        ez[r] = bload(r_ez, o);
        hx[r] = bload(r_hx, o);
        hy[r] = bload(r_hy, o);
      }
      if (POLES & 1) {
        const __amdgpu_buffer_rsrc_t r_jd = plane_rsrc(p.jd_in + base, plane_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) jd[r] = bload(r_jd, row_off(r));
      }
    }
    // this TX's source, if it falls inside the window (halo included)
    const int sc = p.src_cell[bt];
    const int s_row = sc / p.pitch - a0;
    const int s_lane = sc % p.pitch - cb0;
    const bool src_here = s_row >= 0 && s_row < R && s_lane >= 0 && s_lane < 64;

#pragma unroll
    for (int s = 0; s < K; ++s) {
      if (s >= p.ksteps) break;
      const int step = p.step0 + s;
      // H half step (run_cpu form; the stale outer ring reads zeros)
#pragma unroll
      for (int r = s; r < R - s - 1; ++r) {
        const float eR = from_next_lane(ez[r]);
        const float eD = (r + 1 < R) ? ez[r + 1] : 0.f;
        hx[r] = h_minus(hx[r], ch, eR - ez[r]);
        hy[r] = h_plus(hy[r], ch, eD - ez[r]);
      }
      // E step (+ Debye ADE)
#pragma unroll
      for (int r = s + 1; r < R - s - 1; ++r) {
        const float hxl = from_prev_lane(hx[r]);
        const float hyu = (r > 0) ? hy[r - 1] : 0.f;
        const float curl = (hy[r] - hyu) - (hx[r] - hxl);
        const float e0 = ez[r];
        if (POLES == 0) {
          ez[r] = e_step(mca[r], mcb[r], e0, curl, idx);
        } else {
          const float en = e_step_debye(mca[r], mcb[r], e0, curl, idx, mk1[r], jd[r]);
          jd[r] = j_debye(mk1[r], mk2[r], jd[r], en - e0);
          ez[r] = en;
        }
      }
      if (src_here) {
        const float w = p.wave[(long long)bt * p.wave_stride_b + step];
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r == s_row && lane == s_lane) ez[r] += w;
      }
      if (p.e_save) {   // E after this step on the output window (adjoint store)
        const __amdgpu_buffer_rsrc_t r_es =
            plane_rsrc(p.e_save + (long long)s * p.slot + base, plane_bytes);
#pragma unroll
        for (int r = K; r < R - K; ++r) bstore(ez[r], r_es, out_lane ? row_off(r) : kOffNone);
      }
      // probes inside this wavefront's output window
      for (unsigned long long m = pmask; m;) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int qr = __shfl(pr_h, l);
        const int ql = __shfl(pl_h, l);
        float v = 0.f;
#pragma unroll
        for (int r = 0; r < R; ++r) v = (r == qr) ? ez[r] : v;
        if (lane == ql) p.traces[((long long)bt * p.P + l) * p.T + step] = v;
      }
      for (int q0 = 0; !few_probes && q0 < p.P; q0 += 64) {
        const int q = q0 + lane;
        const int pc = q < p.P ? p.probe_cell[q] : -1;
        int pr = -1, pl = -1;
        if (pc >= 0) {
          pr = pc / p.pitch - a0;
          pl = pc % p.pitch - cb0;
        }
        const bool mine = pc >= 0 && pr >= K && pr < R - K && pl >= K && pl < 64 - K;
        unsigned long long m = __ballot(mine);
        while (m) {
          const int l = __ffsll((long long)m) - 1;
          m &= m - 1;
          const int qr = __shfl(pr, l);
          const int ql = __shfl(pl, l);
          float v = 0.f;
#pragma unroll
          for (int r = 0; r < R; ++r) v = (r == qr) ? ez[r] : v;
          if (lane == ql) p.traces[((long long)bt * p.P + q0 + l) * p.T + step] = v;
        }
      }
    }
    // write back the exact inner window
    {
      const __amdgpu_buffer_rsrc_t r_ez = plane_rsrc(p.ez_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hx = plane_rsrc(p.hx_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hy = plane_rsrc(p.hy_out + base, plane_bytes);
#pragma unroll
      for (int r = K; r < R - K; ++r) {
        const unsigned o = out_lane ? row_off(r) : kOffNone;
        bstore(ez[r], r_ez, o);
        bstore(hx[r], r_hx, o);
        bstore(hy[r], r_hy, o);
      }
      if (POLES == 1) {
        const __amdgpu_buffer_rsrc_t r_jd = plane_rsrc(p.jd_out + base, plane_bytes);
#pragma unroll
        for (int r = K; r < R - K; ++r) bstore(jd[r], r_jd, out_lane ? row_off(r) : kOffNone);
      }
    }
  }
}

// Transpose of step2d_tb_kernel's non-dispersive update.  Per adjoint step,
// with q = cbi * lam_e:
//   mu_x += q(c+1) - q(c);  mu_y += q(a) - q(a+1);
//   lam_e = ca lam_e + ch [(mu_y(a-1) - mu_y(a)) - (mu_x(c-1) - mu_x(c))] + dL/dtrace(s-1)
// The window shrinks by one cell per step exactly as in the forward kernel.
template <int R, int K>
__global__ __launch_bounds__(256) void adj2d_tb_kernel(AdjTB2DArgs p) {
  constexpr int NO = R - 2 * K;                        // output rows
  const int lane = threadIdx.x & 63;
  const int task = xcd_block(p.xcd_order) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (task >= p.ntasks) return;
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = task % tiles;
  const int chunk = task / tiles;
  const int ty = tile % p.tiles_y;
  const int tx = tile / p.tiles_y;
  const int a0 = tx * NO - K;
  const int cb0 = ty * (64 - 2 * K) - K;
  const int c = cb0 + lane;
  const bool col_ok = c >= 0 && c < p.My;
  const bool out_lane = lane >= K && lane < 64 - K && col_ok;

  float mca[R], mcbi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int a = a0 + r;
    const bool in = col_ok && a >= 0 && a < p.Mx;
    const long long cell = (long long)a * p.pitch + c;
    mca[r] = in ? p.ca[cell] : 0.f;
    mcbi[r] = in ? p.cb[cell] * p.idx : 0.f;
  }
  float gca[NO], gcb[NO];
#pragma unroll
  for (int r = 0; r < NO; ++r) { gca[r] = 0.f; gcb[r] = 0.f; }
  const float ch = p.ch;
  // probes anywhere in the window (halo included) receive dL/dtrace; lane q
  // describes probe q when P <= 64
  const bool few_probes = p.P <= 64;
  int pr_h = -1, pl_h = -1;
  unsigned long long pmask = 0;
  if (few_probes) {
    const int pc = lane < p.P ? p.probe_cell[lane] : -1;
    if (pc >= 0) {
      pr_h = pc / p.pitch - a0;
      pl_h = pc % p.pitch - cb0;
    }
    pmask = __ballot(pc >= 0 && pr_h >= 0 && pr_h < R && pl_h >= 0 && pl_h < 64);
  }

  const int b_begin = chunk * p.bpw;
  const int b_end = min(b_begin + p.bpw, p.B);
  for (int bt = b_begin; bt < b_end; ++bt) {
    const long long base = (long long)bt * p.bstride;
    float le[R], mx[R], my[R], e1[NO];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int a = a0 + r;
      const bool in = col_ok && a >= 0 && a < p.Mx;
      const long long o = base + (long long)a * p.pitch + c;
      le[r] = in ? p.le_in[o] : 0.f;
      mx[r] = in ? p.mx_in[o] : 0.f;
      my[r] = in ? p.my_in[o] : 0.f;
    }
    const float* e_hi = p.e_store + (long long)(p.step_hi - p.seg0) * p.slot + base;
#pragma unroll
    for (int r = 0; r < NO; ++r) {
      const int a = a0 + K + r;
      e1[r] = (out_lane && a < p.Mx) ? e_hi[(long long)a * p.pitch + c] : 0.f;
    }
    const int sc = p.src_cell[bt];
    const int s_row = sc / p.pitch - a0;
    const int s_lane = sc % p.pitch - cb0;
    // everything the K steps read from memory, issued before the first step
    float es_all[K][NO], wsrc[K], gq[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int step = p.step_hi - 1 - k;
      const bool live = k < p.ksteps;
      const float* e_s = p.e_store + (long long)(step - p.seg0) * p.slot + base;
#pragma unroll
      for (int r = 0; r < NO; ++r) {
        const int a = a0 + K + r;
        es_all[k][r] = (live && out_lane && a < p.Mx) ? e_s[(long long)a * p.pitch + c] : 0.f;
      }
      wsrc[k] = live ? p.wave[(long long)bt * p.wave_stride_b + step] : 0.f;
      gq[k] = (live && step >= 1 && few_probes && ((pmask >> lane) & 1ull))
                  ? p.dtr[((long long)bt * p.P + lane) * p.T + (step - 1)] : 0.f;
    }

#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k >= p.ksteps) break;
      const int step = p.step_hi - 1 - k;
      // gradient sums on the output window (lam_e here is lam_e^{step+1})
#pragma unroll
      for (int r = 0; r < NO; ++r) {
        const float es = es_all[k][r];
        float d = e1[r] - mca[K + r] * es;
        if (K + r == s_row && lane == s_lane) d -= wsrc[k];
        gca[r] += le[K + r] * es;
        gcb[r] += le[K + r] * d;
        e1[r] = es;
      }
      // adjoint H half step
      float q[R];
#pragma unroll
      for (int r = 0; r < R; ++r) q[r] = mcbi[r] * le[r];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float qR = from_next_lane(q[r]);
        const float qD = (r + 1 < R) ? q[r + 1] : 0.f;
        mx[r] = mx[r] + (qR - q[r]);
        my[r] = my[r] + (q[r] - qD);
      }
      // adjoint E step
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float mxL = from_prev_lane(mx[r]);
        const float myU = (r > 0) ? my[r - 1] : 0.f;
        le[r] = mca[r] * le[r] + ch * ((myU - my[r]) - (mxL - mx[r]));
      }
      // dL/dtrace(step - 1) enters lam_e^step at the probes (halo included)
      for (unsigned long long m = pmask; m;) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int qr = __shfl(pr_h, l);
        const int ql = __shfl(pl_h, l);
        const float g = __shfl(gq[k], l);
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r == qr && lane == ql) le[r] += g;
      }
      if (!few_probes && step >= 1) {
        for (int q0 = 0; q0 < p.P; q0 += 64) {
          const int qi = q0 + lane;
          const int pc = qi < p.P ? p.probe_cell[qi] : -1;
          int pr = -1, pl = -1;
          if (pc >= 0) {
            pr = pc / p.pitch - a0;
            pl = pc % p.pitch - cb0;
          }
          const bool here = pc >= 0 && pr >= 0 && pr < R && pl >= 0 && pl < 64;
          unsigned long long m = __ballot(here);
          while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const int qr = __shfl(pr, l);
            const int ql = __shfl(pl, l);
            const float g = p.dtr[((long long)bt * p.P + q0 + l) * p.T + (step - 1)];
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (r == qr && lane == ql) le[r] += g;
          }
        }
      }
    }
#pragma unroll
    for (int r = K; r < R - K; ++r) {
      const int a = a0 + r;
      if (out_lane && a < p.Mx) {
        const long long o = base + (long long)a * p.pitch + c;
        p.le_out[o] = le[r];
        p.mx_out[o] = mx[r];
        p.my_out[o] = my[r];
      }
    }
  }
  // this wavefront is the only writer of its output cells in its chunk's
  // partial-sum plane during this launch: plain read-modify-write
  float* gca_p = p.g_ca + (long long)chunk * p.gstride;
  float* gcb_p = p.g_cbi + (long long)chunk * p.gstride;
#pragma unroll
  for (int r = 0; r < NO; ++r) {
    const int a = a0 + K + r;
    if (out_lane && a < p.Mx && mcbi[K + r] != 0.f) {
      const long long cell = (long long)a * p.pitch + c;
      gca_p[cell] += gca[r];
      gcb_p[cell] += gcb[r];
    }
  }
}

// Reversible form of adj2d_tb_kernel.  Each reversed step first rebuilds
// the forward state one step back from the window's own registers
//   E^s = (E^{s+1} - source(s) - cb/dx curl H^{s+1/2}) / ca      (E interior; 0 elsewhere)
//   hx^{s-1/2} = hx^{s+1/2} + ch (E^s(c+1) - E^s(c)),  hy^{s-1/2} = hy^{s+1/2} - ch (E^s(a+1) - E^s(a))
// (the forward H/E steps of step2d_tb_kernel solved for their inputs), takes
// the sensitivities g_ca += lam E^s and g_cbi += lam cb/dx curl H^{s+1/2} --
// the latter without the cancellation of E^{s+1} - ca E^s - source -- and
// then applies the same adjoint step as adj2d_tb_kernel.  The forward
// reconstruction shrinks the valid window like the forward step (left/top
// for E, right/bottom for H), so the exact inner window is the same.  Every
// segment restarts from an exact checkpoint, which bounds the rounding drift
// of the reconstruction to one segment (the (1/ca)^n growth in lossy cells).
//
// Memory access is through buffer resources: one 32-bit byte offset per
// window row (kOffNone off the box), so the 6 x R loads and the output
// stores carry no per-cell branches and no 64-bit address arithmetic.  As in
// the forward task only rows that reach the output window are updated: at
// reversed step k (0-based) E is rebuilt on rows [k+1, R-k), H on
// [k+1, R-k-1), q is formed on [k, R-k), mu on [k, R-k-1), lam on [k+1, R-k-1).
template <int R, int K>
__global__ __launch_bounds__(256) void adj2d_rev_tb_kernel(AdjTB2DArgs p) {
  constexpr int NO = R - 2 * K;
  const int lane = threadIdx.x & 63;
  const int task = xcd_block(p.xcd_order) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (task >= p.ntasks) return;
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = task % tiles;
  const int chunk = task / tiles;
  const int ty = tile % p.tiles_y;
  const int tx = tile / p.tiles_y;
  const int a0 = tx * NO - K;
  const int cb0 = ty * (64 - 2 * K) - K;
  const int c = cb0 + lane;
  const bool col_ok = c >= 0 && c < p.My;
  const bool out_lane = lane >= K && lane < 64 - K && col_ok;
  // byte offset of this lane's cell in window row r (a0 may be negative: the
  // unsigned sum wraps back to the true offset on the rows inside the box)
  const unsigned pitch4 = (unsigned)p.pitch * 4u;
  const unsigned lane_off = col_ok ? (unsigned)(a0 * p.pitch + c) * 4u : kOffNone;
  const unsigned plane_bytes = (unsigned)p.Mx * pitch4;
  auto row_off = [&](int r) -> unsigned {
    const int a = a0 + r;
    return (a >= 0 && a < p.Mx) ? lane_off + (unsigned)r * pitch4 : kOffNone;
  };

  float mca[R], mica[R], mcbi[R];
  {
    const __amdgpu_buffer_rsrc_t rca = plane_rsrc(p.ca, plane_bytes);
    const __amdgpu_buffer_rsrc_t rcb = plane_rsrc(p.cb, plane_bytes);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      mca[r] = bload(rca, row_off(r));
      mcbi[r] = bload(rcb, row_off(r)) * p.idx;
      mica[r] = mca[r] != 0.f ? 1.f / mca[r] : 0.f;
    }
  }
  float gca[NO], gcb[NO];
#pragma unroll
  for (int r = 0; r < NO; ++r) { gca[r] = 0.f; gcb[r] = 0.f; }
  const float ch = p.ch;
  const bool few_probes = p.P <= 64;
  int pr_h = -1, pl_h = -1;
  unsigned long long pmask = 0;
  if (few_probes) {
    const int pc = lane < p.P ? p.probe_cell[lane] : -1;
    if (pc >= 0) {
      pr_h = pc / p.pitch - a0;
      pl_h = pc % p.pitch - cb0;
    }
    pmask = __ballot(pc >= 0 && pr_h >= 0 && pr_h < R && pl_h >= 0 && pl_h < 64);
  }

  const int b_begin = chunk * p.bpw;
  const int b_end = min(b_begin + p.bpw, p.B);
  for (int bt = b_begin; bt < b_end; ++bt) {
    const long long base = (long long)bt * p.bstride;
    float le[R], mx[R], my[R], fe[R], fx[R], fy[R];
    {
      const __amdgpu_buffer_rsrc_t r_le = plane_rsrc(p.le_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_mx = plane_rsrc(p.mx_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_my = plane_rsrc(p.my_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fe = plane_rsrc(p.fe_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fx = plane_rsrc(p.fhx_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fy = plane_rsrc(p.fhy_in + base, plane_bytes);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const unsigned o = row_off(r);
        le[r] = bload(r_le, o);
        mx[r] = bload(r_mx, o);
        my[r] = bload(r_my, o);
        fe[r] = bload(r_fe, o);
        fx[r] = bload(r_fx, o);
        fy[r] = bload(r_fy, o);
      }
    }
    const int sc = p.src_cell[bt];
    const int s_row = sc / p.pitch - a0;
    const int s_lane = sc % p.pitch - cb0;
    const bool src_rows = s_row >= 0 && s_row < R;   // uniform
    const bool src_lane = lane == s_lane;
    float wsrc[K], gq[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int step = p.step_hi - 1 - k;
      const bool live = k < p.ksteps;
      wsrc[k] = live ? p.wave[(long long)bt * p.wave_stride_b + step] : 0.f;
      gq[k] = (live && step >= 1 && few_probes && ((pmask >> lane) & 1ull))
                  ? p.dtr[((long long)bt * p.P + lane) * p.T + (step - 1)] : 0.f;
    }

#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k >= p.ksteps) break;
      const int step = p.step_hi - 1 - k;
      // E^{s+1} minus the source added after the forward E step
      if (src_rows) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r == s_row) fe[r] -= src_lane ? wsrc[k] : 0.f;
      }
      // forward E one step back (in place: row r's E is only read by row r),
      // and the sensitivities (lam here is lam^{step+1})
#pragma unroll
      for (int r = k + 1; r < R - k; ++r) {
        const float hxl = from_prev_lane(fx[r]);
        const float hyu = (r > 0) ? fy[r - 1] : 0.f;
        const float curl = (fy[r] - hyu) - (fx[r] - hxl);
        const float dterm = mcbi[r] * curl;
        const float es = (fe[r] - dterm) * mica[r];
        if (r >= K && r < R - K) {
          gca[r - K] += le[r] * es;
          gcb[r - K] += le[r] * dterm;
        }
        fe[r] = es;
      }
      // forward H one step back
#pragma unroll
      for (int r = k + 1; r < R - k - 1; ++r) {
        const float eR = from_next_lane(fe[r]);
        const float eD = (r + 1 < R) ? fe[r + 1] : 0.f;
        fx[r] = fx[r] + ch * (eR - fe[r]);
        fy[r] = fy[r] - ch * (eD - fe[r]);
      }
      // adjoint H half step and E step (as adj2d_tb_kernel)
      float q[R];
#pragma unroll
      for (int r = k; r < R - k; ++r) q[r] = mcbi[r] * le[r];
#pragma unroll
      for (int r = k; r < R - k - 1; ++r) {
        const float qR = from_next_lane(q[r]);
        const float qD = (r + 1 < R) ? q[r + 1] : 0.f;
        mx[r] = mx[r] + (qR - q[r]);
        my[r] = my[r] + (q[r] - qD);
      }
#pragma unroll
      for (int r = k + 1; r < R - k - 1; ++r) {
        const float mxL = from_prev_lane(mx[r]);
        const float myU = (r > 0) ? my[r - 1] : 0.f;
        le[r] = mca[r] * le[r] + ch * ((myU - my[r]) - (mxL - mx[r]));
      }
      for (unsigned long long m = pmask; m;) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int qr = __shfl(pr_h, l);
        const int ql = __shfl(pl_h, l);
        const float g = __shfl(gq[k], l);
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r == qr && lane == ql) le[r] += g;
      }
      if (!few_probes && step >= 1) {
        for (int q0 = 0; q0 < p.P; q0 += 64) {
          const int qi = q0 + lane;
          const int pc = qi < p.P ? p.probe_cell[qi] : -1;
          int pr = -1, pl = -1;
          if (pc >= 0) {
            pr = pc / p.pitch - a0;
            pl = pc % p.pitch - cb0;
          }
          const bool here = pc >= 0 && pr >= 0 && pr < R && pl >= 0 && pl < 64;
          unsigned long long m = __ballot(here);
          while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const int qr = __shfl(pr, l);
            const int ql = __shfl(pl, l);
            const float g = p.dtr[((long long)bt * p.P + q0 + l) * p.T + (step - 1)];
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (r == qr && lane == ql) le[r] += g;
          }
        }
      }
    }
    {
      const __amdgpu_buffer_rsrc_t r_le = plane_rsrc(p.le_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_mx = plane_rsrc(p.mx_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_my = plane_rsrc(p.my_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fe = plane_rsrc(p.fe_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fx = plane_rsrc(p.fhx_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fy = plane_rsrc(p.fhy_out + base, plane_bytes);
#pragma unroll
      for (int r = K; r < R - K; ++r) {
        const unsigned o = out_lane ? row_off(r) : kOffNone;
        bstore(le[r], r_le, o);
        bstore(mx[r], r_mx, o);
        bstore(my[r], r_my, o);
        bstore(fe[r], r_fe, o);
        bstore(fx[r], r_fx, o);
        bstore(fy[r], r_fy, o);
      }
    }
  }
  // this wavefront is the only writer of its output cells in its chunk's
  // partial-sum plane during this launch: plain read-modify-write
  const __amdgpu_buffer_rsrc_t rgca = plane_rsrc(p.g_ca + (long long)chunk * p.gstride, plane_bytes);
  const __amdgpu_buffer_rsrc_t rgcb = plane_rsrc(p.g_cbi + (long long)chunk * p.gstride, plane_bytes);
#pragma unroll
  for (int r = 0; r < NO; ++r) {
    const unsigned o = (out_lane && mcbi[K + r] != 0.f) ? row_off(K + r) : kOffNone;
    bstore(bload(rgca, o) + gca[r], rgca, o);
    bstore(bload(rgcb, o) + gcb[r], rgcb, o);
  }
}

template <int R, int K, int POLES>
__global__ __launch_bounds__(256) void step2d_tb_kernel(TB2DArgs p) {
  const int lane = threadIdx.x & 63;
  // readfirstlane: the wavefront index is uniform, and saying so keeps the
  // task decomposition and every derived address in scalar registers
  const int task = xcd_block(p.xcd_order) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (task >= p.ntasks) return;                       // uniform per wavefront
  tb_task<R, K, POLES>(p, task, lane);
}

// Uneven batch, one launch per round: the launch covers that round's
// segments [seg0, seg1), and the launch boundary orders rounds.
template <int R, int K, int POLES>
__global__ __launch_bounds__(256) void round_tb_kernel(MultiArgs m, int seg0, int seg1,
                                                       long long first, long long count) {
  const int lane = threadIdx.x & 63;
  // neighbouring tiles of a problem are consecutive items: with xcd_order
  // each XCD runs a contiguous range of them and shares their halo rows in L2
  const long long w = (long long)xcd_block(m.xcd_order) * 4 +
                      __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (w >= count) return;
  const long long i = first + w;
  int lo = seg0, hi = seg1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.seg_start[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const int pidx = m.seg_prob[lo];
  const int round = m.seg_round[lo];
  const int task = (int)(i - m.seg_start[lo]);
  const MultiDesc& d = m.probs[pidx];
  TB2DArgs a = d.a;
  const int cur = round & 1, nxt = cur ^ 1;
  a.ez_in = d.f[cur][0]; a.hx_in = d.f[cur][1]; a.hy_in = d.f[cur][2];
  a.ez_out = d.f[nxt][0]; a.hx_out = d.f[nxt][1]; a.hy_out = d.f[nxt][2];
  a.jd_in = d.jd[cur]; a.jd_out = d.jd[nxt];
  a.step0 = round * K;
  a.ksteps = min(K, a.T - a.step0);
  a.e_save = nullptr;
  tb_task<R, K, POLES>(a, task, lane);
}

template <int R, int K>
static hipError_t launch_round_rk(const MultiArgs& m, int poles, int seg0, int seg1,
                                  long long first, long long count, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((unsigned)((count + 3) / 4), 1, 1);
  if (poles == 0) hipLaunchKernelGGL((round_tb_kernel<R, K, 0>), grd, blk, 0, s, m, seg0, seg1, first, count);
  else if (poles == 1) hipLaunchKernelGGL((round_tb_kernel<R, K, 1>), grd, blk, 0, s, m, seg0, seg1, first, count);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_round_tb(const MultiArgs& m, int poles, int R, int K, int seg0, int seg1,
                           long long first, long long count, hipStream_t s) {
#define CEEP_RR(RR)                                                                        \
  switch (K) {                                                                             \
    case 1: return launch_round_rk<RR, 1>(m, poles, seg0, seg1, first, count, s);           \
    case 2: return launch_round_rk<RR, 2>(m, poles, seg0, seg1, first, count, s);           \
    case 3: return launch_round_rk<RR, 3>(m, poles, seg0, seg1, first, count, s);           \
    case 4: return launch_round_rk<RR, 4>(m, poles, seg0, seg1, first, count, s);           \
  }                                                                                        \
  return hipErrorInvalidValue;
  if (R == 16) { CEEP_RR(16) }
  if (R == 24) { CEEP_RR(24) }
  if (R == 32) { CEEP_RR(32) }
#undef CEEP_RR
  return hipErrorInvalidValue;
}

template <int R, int K>
static hipError_t launch_rk(const TB2DArgs& a, int poles, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((a.ntasks + 3) / 4, 1, 1);
  if (poles == 0) hipLaunchKernelGGL((step2d_tb_kernel<R, K, 0>), grd, blk, 0, s, a);
  else if (poles == 1) hipLaunchKernelGGL((step2d_tb_kernel<R, K, 1>), grd, blk, 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int R>
static hipError_t launch_r(const TB2DArgs& a, int poles, int K, hipStream_t s) {
  switch (K) {
    case 1: return launch_rk<R, 1>(a, poles, s);
    case 2: return launch_rk<R, 2>(a, poles, s);
    case 3: return launch_rk<R, 3>(a, poles, s);
    case 4: return launch_rk<R, 4>(a, poles, s);
  }
  return hipErrorInvalidValue;
}

template <int R>
static hipError_t launch_adj_r(const AdjTB2DArgs& a, int K, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((a.ntasks + 3) / 4, 1, 1);
  switch (K) {
    case 1: hipLaunchKernelGGL((adj2d_tb_kernel<R, 1>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((adj2d_tb_kernel<R, 2>), grd, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((adj2d_tb_kernel<R, 3>), grd, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((adj2d_tb_kernel<R, 4>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int R>
static hipError_t launch_rev_r(const AdjTB2DArgs& a, int K, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((a.ntasks + 3) / 4, 1, 1);
  switch (K) {
    case 1: hipLaunchKernelGGL((adj2d_rev_tb_kernel<R, 1>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((adj2d_rev_tb_kernel<R, 2>), grd, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((adj2d_rev_tb_kernel<R, 3>), grd, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((adj2d_rev_tb_kernel<R, 4>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_adj2d_rev_tb(const AdjTB2DArgs& a, int R, int K, hipStream_t s) {
  switch (R) {
    case 16: return launch_rev_r<16>(a, K, s);
    case 24: return launch_rev_r<24>(a, K, s);
    case 32: return launch_rev_r<32>(a, K, s);
  }
  return hipErrorInvalidValue;
}

// Multi-wavefront form of adj2d_rev_tb_kernel: one workgroup owns an
// R x 64 window (R = NW RW, NW = 4 or 8 wavefronts) and each of its
// wavefronts RW consecutive rows.  Each reversed step needs one row across
// every wavefront boundary four times -- the old hy above (forward E), the
// new E below (forward H), q below (adjoint H) and the new mu_y above (lam)
// -- and passes them through LDS: E and q at one workgroup barrier, mu_y
// together with the hy of the next reversed step at a second, so a launch
// of K reversed steps per TX takes 2K + 1 barriers.  The
// arithmetic per cell is that of adj2d_rev_tb_kernel; the results agree to
// FP32 rounding (instantiations may fuse different multiply-adds).  A wavefront's registers grow by about 17
// VGPRs per row (six fields, three maps, two gradient sums and the shifted
// neighbours): 92 at RW = 4, 128 at RW = 6 (with the occupancy bound below),
// 162 at RW = 8.  Eight wavefronts of 6 rows therefore hold 48-row windows
// (61 % of the window exact at K = 6, against 51 % for 32 rows) at four
// wavefronts per SIMD, where four of 8 rows hold 32-row windows at three.
//
template <int RW, int NW>
constexpr int adj_mw_min_waves() {
  return (NW == 8 && RW == 6) ? 4 : 1;
}

// PF: the six planes of the chunk's next TX are loaded into a second register
// set while the current TX is stepped (its dL/dtrace samples are issued
// first, since the in-order vector-memory counter would otherwise make their
// wait drain the prefetch); the set moves over at the next TX.  About 24
// more VGPRs at RW = 4, 48 at RW = 8; results are bit-identical.
template <int RW, int K, int NW, bool PF = false>
__global__ __launch_bounds__(64 * NW, (adj_mw_min_waves<RW, NW>())) void adj2d_rev_mw_kernel(
    AdjTB2DArgs p) {
  constexpr int R = NW * RW;
  constexpr int NO = R - 2 * K;
  // exchange rows with one zero row beyond the window's edge, as in mw_task:
  // rows read from the wavefront above (x_fy, x_my) are written at w + 1 and
  // read at w (row 0 zero), rows from the one below (x_es, x_q) written at w
  // and read at w + 1 (row NW zero), so the edge wavefronts read no branch
  __shared__ float x_fy[NW + 1][64], x_es[NW + 1][64], x_q[NW + 1][64], x_my[NW + 1][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int task = xcd_block(p.xcd_order);
  if (task >= p.ntasks) return;                       // uniform per workgroup
  if (w == 0) {                                       // before the first barrier
    x_fy[0][lane] = 0.f;
    x_my[0][lane] = 0.f;
    x_es[NW][lane] = 0.f;
    x_q[NW][lane] = 0.f;
  }
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = task % tiles;
  const int chunk = task / tiles;
  const int ty = tile % p.tiles_y;
  const int tx = tile / p.tiles_y;
  const int a0 = tx * NO - K;                          // row of window row 0
  const int cb0 = ty * (64 - 2 * K) - K;
  const int c = cb0 + lane;
  const bool col_ok = c >= 0 && c < p.My;
  const bool out_lane = lane >= K && lane < 64 - K && col_ok;
  const int r0 = w * RW;                               // this wavefront's first window row
  const unsigned pitch4 = (unsigned)p.pitch * 4u;
  const unsigned lane_off = col_ok ? (unsigned)((a0 + r0) * p.pitch + c) * 4u : kOffNone;
  const unsigned plane_bytes = (unsigned)p.Mx * pitch4;
  auto row_off = [&](int i) -> unsigned {               // local row i of this wavefront
    const int a = a0 + r0 + i;
    return (a >= 0 && a < p.Mx) ? lane_off + (unsigned)i * pitch4 : kOffNone;
  };
  auto is_out = [&](int i) -> bool {                    // window row inside [K, R - K)
    return r0 + i >= K && r0 + i < R - K;
  };

  float mca[RW], mica[RW], mcbi[RW];
  {
    const __amdgpu_buffer_rsrc_t rca = plane_rsrc(p.ca, plane_bytes);
    const __amdgpu_buffer_rsrc_t rcb = plane_rsrc(p.cb, plane_bytes);
#pragma unroll
    for (int i = 0; i < RW; ++i) {
      mca[i] = bload(rca, row_off(i));
      mcbi[i] = bload(rcb, row_off(i)) * p.idx;
      mica[i] = mca[i] != 0.f ? 1.f / mca[i] : 0.f;
    }
  }
  float gca[RW], gcb[RW];
#pragma unroll
  for (int i = 0; i < RW; ++i) { gca[i] = 0.f; gcb[i] = 0.f; }
  const float ch = p.ch;
  const bool few_probes = p.P <= 64;
  int pr_h = -1, pl_h = -1;
  unsigned long long pmask = 0;
  if (few_probes) {
    const int pc = lane < p.P ? p.probe_cell[lane] : -1;
    if (pc >= 0) {
      pr_h = pc / p.pitch - (a0 + r0);                 // local row in this wavefront
      pl_h = pc % p.pitch - cb0;
    }
    pmask = __ballot(pc >= 0 && pr_h >= 0 && pr_h < RW && pl_h >= 0 && pl_h < 64);
  }

  const int b_begin = chunk * p.bpw;
  const int b_end = min(b_begin + p.bpw, p.B);
  float le[RW], mx[RW], my[RW], fe[RW], fx[RW], fy[RW];
  constexpr int NP = PF ? RW : 1;
  float nle[NP], nmx[NP], nmy[NP], nfe[NP], nfx[NP], nfy[NP];
  auto load6 = [&](int bt, float* dle, float* dmx, float* dmy, float* dfe, float* dfx,
                   float* dfy) {
    const long long base = (long long)bt * p.bstride;
    const __amdgpu_buffer_rsrc_t r_le = plane_rsrc(p.le_in + base, plane_bytes);
    const __amdgpu_buffer_rsrc_t r_mx = plane_rsrc(p.mx_in + base, plane_bytes);
    const __amdgpu_buffer_rsrc_t r_my = plane_rsrc(p.my_in + base, plane_bytes);
    const __amdgpu_buffer_rsrc_t r_fe = plane_rsrc(p.fe_in + base, plane_bytes);
    const __amdgpu_buffer_rsrc_t r_fx = plane_rsrc(p.fhx_in + base, plane_bytes);
    const __amdgpu_buffer_rsrc_t r_fy = plane_rsrc(p.fhy_in + base, plane_bytes);
#pragma unroll
    for (int i = 0; i < RW; ++i) {
      const unsigned o = row_off(i);
      dle[i] = bload(r_le, o);
      dmx[i] = bload(r_mx, o);
      dmy[i] = bload(r_my, o);
      dfe[i] = bload(r_fe, o);
      dfx[i] = bload(r_fx, o);
      dfy[i] = bload(r_fy, o);
    }
  };
  if (PF && b_begin < b_end) load6(b_begin, le, mx, my, fe, fx, fy);
  for (int bt = b_begin; bt < b_end; ++bt) {
    const long long base = (long long)bt * p.bstride;
    if (!PF) {
      load6(bt, le, mx, my, fe, fx, fy);
    } else if (bt > b_begin) {
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        le[i] = nle[i]; mx[i] = nmx[i]; my[i] = nmy[i];
        fe[i] = nfe[i]; fx[i] = nfx[i]; fy[i] = nfy[i];
      }
    }
    const int sc = p.src_cell[bt];
    const int s_row = sc / p.pitch - (a0 + r0);          // local row
    const int s_lane = sc % p.pitch - cb0;
    const bool src_rows = s_row >= 0 && s_row < RW;      // uniform
    const bool src_lane = lane == s_lane;
    float wsrc[K], gq[K];
    // the waveform sample is uniform: a scalar load, outside the vector-memory count
    typedef __attribute__((address_space(4))) const float cfloat;
    cfloat* wv = (cfloat*)(p.wave + (long long)bt * p.wave_stride_b);
    const bool pq = few_probes && ((pmask >> lane) & 1ull);   // this lane's probe is in the window
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int step = p.step_hi - 1 - k;
      const bool live = k < p.ksteps;
      wsrc[k] = live ? wv[step] : 0.f;
      gq[k] = (live && step >= 1 && pq)
                  ? p.dtr[((long long)bt * p.P + lane) * p.T + (step - 1)] : 0.f;
    }
    if (PF && bt + 1 < b_end) load6(bt + 1, nle, nmx, nmy, nfe, nfx, nfy);

    // forward E one step back needs the forward hy of the row above this
    // wavefront's first row, from the wavefront above: for the first reversed
    // step here, for later ones with mu_y at the step before's second barrier
    x_fy[w + 1][lane] = fy[RW - 1];
    lds_barrier();
    float fy_up = x_fy[w][lane];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k >= p.ksteps) break;                        // uniform per workgroup
      const int step = p.step_hi - 1 - k;
      const bool more = k + 1 < p.ksteps;              // uniform
      if (src_rows) {
#pragma unroll
        for (int i = 0; i < RW; ++i)
          if (i == s_row) fe[i] -= src_lane ? wsrc[k] : 0.f;
      }
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const float hxl = from_prev_lane(fx[i]);
        const float hyu = (i > 0) ? fy[i - 1] : fy_up;
        const float curl = (fy[i] - hyu) - (fx[i] - hxl);
        const float dterm = mcbi[i] * curl;
        const float es = (fe[i] - dterm) * mica[i];
        gca[i] += le[i] * es;
        gcb[i] += le[i] * dterm;
        fe[i] = es;
      }
      // forward H one step back and the adjoint H step: the rows below come
      // from the wavefront below (its new E and its q)
      x_es[w][lane] = fe[0];
      x_q[w][lane] = mcbi[0] * le[0];
      lds_barrier();
      const float es_dn = x_es[w + 1][lane];
      const float q_dn = x_q[w + 1][lane];
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const float eR = from_next_lane(fe[i]);
        const float eD = (i + 1 < RW) ? fe[i + 1] : es_dn;
        fx[i] = fx[i] + ch * (eR - fe[i]);
        fy[i] = fy[i] - ch * (eD - fe[i]);
      }
      float q[RW];
#pragma unroll
      for (int i = 0; i < RW; ++i) q[i] = mcbi[i] * le[i];
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const float qR = from_next_lane(q[i]);
        const float qD = (i + 1 < RW) ? q[i + 1] : q_dn;
        mx[i] = mx[i] + (qR - q[i]);
        my[i] = my[i] + (q[i] - qD);
      }
      // adjoint E step: the new mu_y above comes from the wavefront above,
      // and with it the forward hy the next reversed step starts from
      x_my[w + 1][lane] = my[RW - 1];
      if (more) x_fy[w + 1][lane] = fy[RW - 1];
      lds_barrier();
      const float my_up = x_my[w][lane];
      if (more) fy_up = x_fy[w][lane];
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const float mxL = from_prev_lane(mx[i]);
        const float myU = (i > 0) ? my[i - 1] : my_up;
        le[i] = mca[i] * le[i] + ch * ((myU - my[i]) - (mxL - mx[i]));
      }
      for (unsigned long long m = pmask; m;) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int qr = __shfl(pr_h, l);
        const int ql = __shfl(pl_h, l);
        const float g = __shfl(gq[k], l);
#pragma unroll
        for (int i = 0; i < RW; ++i)
          if (i == qr && lane == ql) le[i] += g;
      }
      if (!few_probes && step >= 1) {
        for (int q0 = 0; q0 < p.P; q0 += 64) {
          const int qi = q0 + lane;
          const int pc = qi < p.P ? p.probe_cell[qi] : -1;
          int pr = -1, pl = -1;
          if (pc >= 0) {
            pr = pc / p.pitch - (a0 + r0);
            pl = pc % p.pitch - cb0;
          }
          const bool here = pc >= 0 && pr >= 0 && pr < RW && pl >= 0 && pl < 64;
          unsigned long long m = __ballot(here);
          while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const int qr = __shfl(pr, l);
            const int ql = __shfl(pl, l);
            const float g = p.dtr[((long long)bt * p.P + q0 + l) * p.T + (step - 1)];
#pragma unroll
            for (int i = 0; i < RW; ++i)
              if (i == qr && lane == ql) le[i] += g;
          }
        }
      }
    }
    {
      const __amdgpu_buffer_rsrc_t r_le = plane_rsrc(p.le_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_mx = plane_rsrc(p.mx_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_my = plane_rsrc(p.my_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fe = plane_rsrc(p.fe_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fx = plane_rsrc(p.fhx_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_fy = plane_rsrc(p.fhy_out + base, plane_bytes);
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const unsigned o = (out_lane && is_out(i)) ? row_off(i) : kOffNone;
        bstore(le[i], r_le, o);
        bstore(mx[i], r_mx, o);
        bstore(my[i], r_my, o);
        bstore(fe[i], r_fe, o);
        bstore(fx[i], r_fx, o);
        bstore(fy[i], r_fy, o);
      }
    }
  }
  // the workgroup is the only writer of its output cells in its chunk's
  // partial-sum plane during this launch: plain read-modify-write
  const __amdgpu_buffer_rsrc_t rgca = plane_rsrc(p.g_ca + (long long)chunk * p.gstride, plane_bytes);
  const __amdgpu_buffer_rsrc_t rgcb = plane_rsrc(p.g_cbi + (long long)chunk * p.gstride, plane_bytes);
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    const unsigned o = (out_lane && is_out(i) && mcbi[i] != 0.f) ? row_off(i) : kOffNone;
    bstore(bload(rgca, o) + gca[i], rgca, o);
    bstore(bload(rgcb, o) + gcb[i], rgcb, o);
  }
}

template <int RW>
static hipError_t launch_rev_mw_r(const AdjTB2DArgs& a, int K, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((unsigned)a.ntasks, 1, 1);
  switch (K) {
    case 1: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 1, 4>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 2, 4>), grd, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 3, 4>), grd, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 4, 4>), grd, blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 5, 4>), grd, blk, 0, s, a); break;
    case 6:
      if (RW < 8) return hipErrorInvalidValue;   // the window keeps at least 20 exact rows
      hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 6, 4>), grd, blk, 0, s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// eight wavefronts: 32-row windows of 4 rows per wavefront (K <= 6) or
// 48-row windows of 6 (K = 4..6); every launch depth below K is there too,
// since a segment's last launch may take fewer steps only through ksteps
template <int RW, bool PF>
static hipError_t launch_rev_mw8_r(const AdjTB2DArgs& a, int K, hipStream_t s) {
  const dim3 blk(512, 1, 1), grd((unsigned)a.ntasks, 1, 1);
  switch (K) {
    case 4: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 4, 8, PF>), grd, blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 5, 8, PF>), grd, blk, 0, s, a); break;
    case 6: hipLaunchKernelGGL((adj2d_rev_mw_kernel<RW, 6, 8, PF>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// four wavefronts of 8 rows with the next TX prefetched (depth 6 only)
static hipError_t launch_rev_mw4_pf(const AdjTB2DArgs& a, hipStream_t s) {
  const dim3 blk(256, 1, 1), grd((unsigned)a.ntasks, 1, 1);
  hipLaunchKernelGGL((adj2d_rev_mw_kernel<8, 6, 4, true>), grd, blk, 0, s, a);
  return hipGetLastError();
}

bool adj2d_rev_mw_shape_ok(int R, int K, int NW, bool prefetch) {
  if (NW == 4 && prefetch) return R == 32 && K == 6;
  if (NW == 4) return (R == 16 || R == 24 || R == 32) && K >= 1 && K <= (R == 32 ? 6 : 5);
  if (NW == 8) return (R == 32 || (R == 48 && !prefetch)) && K >= 4 && K <= 6;
  return false;
}

hipError_t launch_adj2d_rev_mw(const AdjTB2DArgs& a, int R, int K, int NW, bool prefetch,
                               hipStream_t s) {
  if (a.ntasks < 1) return hipSuccess;
  if (!adj2d_rev_mw_shape_ok(R, K, NW, prefetch)) return hipErrorInvalidValue;
  if (NW == 8 && prefetch) return launch_rev_mw8_r<4, true>(a, K, s);
  if (NW == 8) return R == 32 ? launch_rev_mw8_r<4, false>(a, K, s) : launch_rev_mw8_r<6, false>(a, K, s);
  if (prefetch) return launch_rev_mw4_pf(a, s);
  switch (R) {
    case 16: return launch_rev_mw_r<4>(a, K, s);
    case 24: return launch_rev_mw_r<6>(a, K, s);
    case 32: return launch_rev_mw_r<8>(a, K, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_adj2d_tb(const AdjTB2DArgs& a, int R, int K, hipStream_t s) {
  switch (R) {
    case 16: return launch_adj_r<16>(a, K, s);
    case 24: return launch_adj_r<24>(a, K, s);
    case 32: return launch_adj_r<32>(a, K, s);
  }
  return hipErrorInvalidValue;
}

// Multi-wavefront form of step2d_tb_kernel.  One workgroup owns an R x 64
// window (R = NW RW, NW = 4 or 8 wavefronts) and each of its wavefronts RW
// consecutive rows.  A step needs one row across each wavefront boundary
// twice -- the old E below (H update) and the new hy above (E update) --
// passed through LDS between two workgroup barriers.  With RW rows per
// wavefront the window can be tall (R = 48 or 64) and the blocking deep (K up
// to 8) within the register file: the loads per output cell-step drop from
// (R/(R-2K) (64/(64-2K)) + 1)/K = 0.82 at R = 16, K = 4 to 0.38 at R = 48,
// K = 8, and the exactly computed fraction of the window at K = 6 rises from
// (20/32)(52/64) = 51 % at R = 32 to 66 % at R = 64.  Eight wavefronts of 8
// rows give 64-row windows at the register footprint of four of 8.
// Per-cell arithmetic is that of step2d_tb_kernel.
// One task (output tile, TX chunk) advanced p.ksteps steps by the NW
// wavefronts of the calling workgroup (or of its group: the task's two
// workgroup barriers per step then hold every group); x_e / x_h are its LDS
// exchange rows, NW + 1 of each: wavefront w writes x_e[w] and x_h[w + 1]
// and reads x_e[w + 1] and x_h[w], so row NW of x_e and row 0 of x_h stand
// for the zero rows beyond the window's edges.  The caller zeroes those two
// rows once, before the task's first barrier; no task writes them.  The
// reads then need no branch, and each lane shift stays in the block of the
// subtraction that uses it, which folds the two into one instruction.
// FULL: the launch takes all K steps (p.ksteps == K).  The step loop then
// has no early exit, so the unrolled steps need not return the fields to
// the registers an exit would read them from (two moves per row and step).
template <int RW, int K, int POLES, bool MID, int NW = 4, bool FULL = false, bool WPRE = false>
__device__ __forceinline__ void mw_task(const TB2DArgs& p, int task, int lane, int w,
                                        float (*x_e)[64], float (*x_h)[64]) {
  constexpr int R = NW * RW;
  const int tiles = p.tile_list ? p.n_list : p.tiles_x * p.tiles_y;
  const int chunk = task / tiles;
  const int tile = p.tile_list ? p.tile_list[task - chunk * tiles] : task - chunk * tiles;
  const int ty = tile % p.tiles_y;
  const int tx = tile / p.tiles_y;
  const int a0 = tx * (R - 2 * K) - K;                // row of window row 0
  const int cb0 = ty * (64 - 2 * K) - K;              // column of lane 0
  const int c = cb0 + lane;
  const bool col_ok = c >= 0 && c < p.My;
  const int r0 = w * RW;                              // this wavefront's first window row
  const unsigned pitch4 = (unsigned)p.pitch * 4u;
  const unsigned lane_off = col_ok ? (unsigned)((a0 + r0) * p.pitch + c) * 4u : kOffNone;
  const unsigned plane_bytes = (unsigned)p.Mx * pitch4;
  auto row_off = [&](int i) -> unsigned {
    const int a = a0 + r0 + i;
    return (a >= 0 && a < p.Mx) ? lane_off + (unsigned)i * pitch4 : kOffNone;
  };
  auto is_out = [&](int i) -> bool { return r0 + i >= K && r0 + i < R - K; };

  float mca[RW], mcb[RW], mk1[RW], mk2[RW];
  if (MID) {
    // piecewise-constant media: one byte per cell (cells off the box read
    // id 0, the all-zero entry), then the (ca, cb, k1, k2dt) table entry
    const __amdgpu_buffer_rsrc_t rid = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<unsigned char*>(p.mat_id), 0, (int)(plane_bytes >> 2), 0x00020000);
    int id[RW];
#pragma unroll
    for (int i = 0; i < RW; ++i)
      id[i] = (int)__builtin_amdgcn_raw_buffer_load_b8(rid, (int)(row_off(i) >> 2), 0, 0);
#pragma unroll
    for (int i = 0; i < RW; ++i) {
      const float4 m = p.mat_table[id[i]];
      mca[i] = m.x;
      mcb[i] = m.y;
      mk1[i] = m.z;
      mk2[i] = m.w;
    }
  } else {
    const __amdgpu_buffer_rsrc_t rca = plane_rsrc(p.ca, plane_bytes);
    const __amdgpu_buffer_rsrc_t rcb = plane_rsrc(p.cb, plane_bytes);
#pragma unroll
    for (int i = 0; i < RW; ++i) {
      mca[i] = bload(rca, row_off(i));
      mcb[i] = bload(rcb, row_off(i));
    }
    if (POLES == 1) {
      const __amdgpu_buffer_rsrc_t rk1 = plane_rsrc(p.d_k1, plane_bytes);
      const __amdgpu_buffer_rsrc_t rk2 = plane_rsrc(p.d_k2dt, plane_bytes);
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        mk1[i] = bload(rk1, row_off(i));
        mk2[i] = bload(rk2, row_off(i));
      }
    }
  }
  // difference form: this tile's window meets cells where the two media
  // differ (uniform per workgroup); mk1 / mk2 then hold d_a / d_b
  const bool dif = POLES == kPolesDiff && p.diff_tile[tile] != 0;
  if (dif) {
    const __amdgpu_buffer_rsrc_t rda = plane_rsrc(p.d_a, plane_bytes);
    const __amdgpu_buffer_rsrc_t rdb = plane_rsrc(p.d_b, plane_bytes);
#pragma unroll
    for (int i = 0; i < RW; ++i) {
      mk1[i] = bload(rda, row_off(i));
      mk2[i] = bload(rdb, row_off(i));
    }
  }
  const float ch = p.ch, idx = p.idx;
  const bool out_lane = lane >= K && lane < 64 - K && col_ok;
  const bool few_probes = p.P <= 64;
  int pr_h = -1, pl_h = -1;
  unsigned long long pmask = 0;
  if (few_probes) {
    const int pc = lane < p.P ? p.probe_cell[lane] : -1;
    int pr_w = -1;
    if (pc >= 0) {
      pr_w = pc / p.pitch - a0;                       // window row
      pr_h = pr_w - r0;                               // this wavefront's local row
      pl_h = pc % p.pitch - cb0;
    }
    pmask = __ballot(pc >= 0 && pr_w >= K && pr_w < R - K && pr_h >= 0 && pr_h < RW &&
                     pl_h >= K && pl_h < 64 - K);
  }

  const int b_begin = chunk * p.bpw;
  const int b_end = min(b_begin + p.bpw, p.B);
  for (int bt = b_begin; bt < b_end; ++bt) {
    const long long base = (long long)bt * p.bstride;
    float ez[RW], hx[RW], hy[RW], jd[RW];
    {
      const __amdgpu_buffer_rsrc_t r_ez = plane_rsrc(p.ez_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hx = plane_rsrc(p.hx_in + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hy = plane_rsrc(p.hy_in + base, plane_bytes);
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const unsigned o = row_off(i);
        ez[i] = bload(r_ez, o);
        hx[i] = bload(r_hx, o);
        hy[i] = bload(r_hy, o);
      }
      if (POLES == 1) {
        const __amdgpu_buffer_rsrc_t r_jd = plane_rsrc(p.jd_in + base, plane_bytes);
#pragma unroll
        for (int i = 0; i < RW; ++i) jd[i] = bload(r_jd, row_off(i));
      }
    }
    // difference form: the reference run's E at the current and the next
    // step (er0, er1), loaded only on cells where the media differ (the
    // source term is 0 elsewhere); each step fetches the one after next
    float er0[RW], er1[RW];
    auto load_ref = [&](float* dst, int step) {
      const __amdgpu_buffer_rsrc_t r_er =
          plane_rsrc(p.e_ref + (long long)step * p.ref_slot + base, plane_bytes);
#pragma unroll
      for (int i = 0; i < RW; ++i)
        dst[i] = bload(r_er, (mk1[i] != 0.f || mk2[i] != 0.f) ? row_off(i) : kOffNone);
    };
    if (dif) {
      load_ref(er0, p.step0);
      load_ref(er1, p.step0 + 1);
    }
    const int sc = p.src_cell[bt];
    const int s_row = sc / p.pitch - (a0 + r0);        // local row
    const int s_lane = sc % p.pitch - cb0;
    const bool src_here = !p.no_inject && s_row >= 0 && s_row < RW && s_lane >= 0 && s_lane < 64;
    // WPRE: the launch's K waveform samples are read up front and
    // unconditionally, as uniform scalar loads (every run holds T samples per
    // TX; the index is clamped for the steps past T of a short last launch,
    // never used).  The sample load under `inject` inside the step splits the
    // block, and the H update's lane shifts, left in the block before the
    // split, then stay separate DPP moves (one per row and step) instead of
    // folding into their subtractions.  Measured on MI355X: the uneven
    // batch's round kernel gains 8 % with WPRE (678 -> 736 GCell-steps/s),
    // while the uniform-batch kernel loses 4 % (842 -> 808), so only the
    // round kernel takes it.
    typedef __attribute__((address_space(4))) const float cfloat;
    cfloat* wvp = (cfloat*)(p.wave + (long long)bt * p.wave_stride_b);
    float wsrc[K];
    if (WPRE) {
#pragma unroll
      for (int k = 0; k < K; ++k) wsrc[k] = wvp[FULL ? p.step0 + k : min(p.step0 + k, p.T - 1)];
    }

#pragma unroll
    for (int s = 0; s < K; ++s) {
      if (!FULL && s >= p.ksteps) break;               // uniform per workgroup
      const int step = p.step0 + s;
      float er2[RW];                                   // difference form: E_ref^{step+2}
      if (dif && s + 1 < (FULL ? K : p.ksteps)) load_ref(er2, step + 2);
      // H half step: the E row below the last row comes from the wavefront below
      x_e[w][lane] = ez[0];
      __syncthreads();
      const float e_dn = x_e[w + 1][lane];             // row NW: zeros
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const float eR = from_next_lane(ez[i]);
        const float eD = (i + 1 < RW) ? ez[i + 1] : e_dn;
        hx[i] = h_minus(hx[i], ch, eR - ez[i]);
        hy[i] = h_plus(hy[i], ch, eD - ez[i]);
      }
      // E step (+ Debye ADE): the new hy above the first row comes from the wavefront above
      x_h[w + 1][lane] = hy[RW - 1];
      __syncthreads();
      const float h_up = x_h[w][lane];                 // row 0: zeros
      // the source (not in the difference form, where it cancels): s_row is
      // uniform, so only that row's update takes the add, on its source lane
      // (a select over every row after the loop would rename all of them and
      // cost two moves per row and step)
      const bool inject = POLES != kPolesDiff && src_here;
      const float wv = !inject ? 0.f
                       : WPRE ? wsrc[s] : p.wave[(long long)bt * p.wave_stride_b + step];
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const float hxl = from_prev_lane(hx[i]);
        const float hyu = (i > 0) ? hy[i - 1] : h_up;
        const float curl = (hy[i] - hyu) - (hx[i] - hxl);
        const float e0 = ez[i];
        if (POLES == 0 || POLES == kPolesDiff) {
          ez[i] = e_step(mca[i], mcb[i], e0, curl, idx);
        } else {
          const float en = e_step_debye(mca[i], mcb[i], e0, curl, idx, mk1[i], jd[i]);
          jd[i] = j_debye(mk1[i], mk2[i], jd[i], en - e0);
          ez[i] = en;
        }
        if (inject && i == s_row && lane == s_lane) ez[i] += wv;
      }
      if (POLES == kPolesDiff) {
        // the source is common to both runs: it cancels in E - E_ref, and
        // E_ref^{s+1} - source^s is the reference's ca_r E_ref^s + cb_r curl
        if (dif) {
          const float wv = !src_here ? 0.f
                           : WPRE ? wsrc[s] : p.wave[(long long)bt * p.wave_stride_b + step];
#pragma unroll
          for (int i = 0; i < RW; ++i) {
            const float e1 = (src_here && i == s_row && lane == s_lane) ? er1[i] - wv : er1[i];
            ez[i] += __builtin_fmaf(mk1[i], er0[i], mk2[i] * e1);
            er0[i] = er1[i];
            er1[i] = er2[i];
          }
        }
      }
      if (p.e_save) {   // E after this step on the output window (adjoint store)
        const __amdgpu_buffer_rsrc_t r_es =
            plane_rsrc(p.e_save + (long long)s * p.slot + base, plane_bytes);
#pragma unroll
        for (int i = 0; i < RW; ++i)
          bstore(ez[i], r_es, (out_lane && is_out(i)) ? row_off(i) : kOffNone);
      }
      for (unsigned long long m = pmask; m;) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int qr = __shfl(pr_h, l);
        const int ql = __shfl(pl_h, l);
        float v = 0.f;
#pragma unroll
        for (int i = 0; i < RW; ++i) v = (i == qr) ? ez[i] : v;
        if (lane == ql) p.traces[((long long)bt * p.P + l) * p.T + step] = v;
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
          float v = 0.f;
#pragma unroll
          for (int i = 0; i < RW; ++i) v = (i == qr) ? ez[i] : v;
          if (lane == ql) p.traces[((long long)bt * p.P + q0 + l) * p.T + step] = v;
        }
      }
    }
    {
      const __amdgpu_buffer_rsrc_t r_ez = plane_rsrc(p.ez_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hx = plane_rsrc(p.hx_out + base, plane_bytes);
      const __amdgpu_buffer_rsrc_t r_hy = plane_rsrc(p.hy_out + base, plane_bytes);
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const unsigned o = (out_lane && is_out(i)) ? row_off(i) : kOffNone;
        bstore(ez[i], r_ez, o);
        bstore(hx[i], r_hx, o);
        bstore(hy[i], r_hy, o);
      }
      if (POLES == 1) {
        const __amdgpu_buffer_rsrc_t r_jd = plane_rsrc(p.jd_out + base, plane_bytes);
#pragma unroll
        for (int i = 0; i < RW; ++i)
          bstore(jd[i], r_jd, (out_lane && is_out(i)) ? row_off(i) : kOffNone);
      }
    }
  }
}

template <int RW, int K, int POLES, bool MID, int NW, bool FULL>
__global__ __launch_bounds__(64 * NW) void step2d_mw_kernel(TB2DArgs p) {
  __shared__ float x_e[NW + 1][64], x_h[NW + 1][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int task = xcd_block(p.xcd_order);
  if (task >= p.ntasks) return;                       // uniform per workgroup
  if (w == 0) {                                       // the edge rows (see mw_task)
    x_e[NW][lane] = 0.f;
    x_h[0][lane] = 0.f;
  }
  mw_task<RW, K, POLES, MID, NW, FULL>(p, task, lane, w, x_e, x_h);
}

// Uneven batch, one launch per round, NW = 4 or 8 wavefronts per window:
// each workgroup is one work item (problem, tile, TX chunk) of the round.
// Eight wavefronts per window: four per SIMD (at most 128 VGPRs), as the
// uniform-batch kernel gets by itself.  Without the bound the Debye form
// takes 134 VGPRs, three wavefronts per SIMD, which leaves one eight-wavefront
// workgroup per CU (64 heads: 400 ms per batch at K = 6, 338 with the bound,
// which spills 5-6 VGPRs).  Four-wavefront workgroups fit three per CU at
// 134 VGPRs; the same bound there made them slower (285 -> 357 ms).
template <int NW>
constexpr int round_mw_min_waves() {
  return NW == 8 ? 4 : 1;
}

template <int RW, int K, int POLES, bool MID, int NW = 4>
__global__ __launch_bounds__(64 * NW, (round_mw_min_waves<NW>())) void round_mw_kernel(
    MultiArgs m, int seg0, int seg1, long long first, long long count) {
  __shared__ float x_e[NW + 1][64], x_h[NW + 1][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long wi = xcd_block(m.xcd_order);
  if (wi >= count) return;                            // uniform per workgroup
  if (w == 0) {                                       // the edge rows (see mw_task)
    x_e[NW][lane] = 0.f;
    x_h[0][lane] = 0.f;
  }
  const long long i = first + wi;
  int lo = seg0, hi = seg1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.seg_start[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const int pidx = m.seg_prob[lo];
  const int round = m.seg_round[lo];
  const int task = (int)(i - m.seg_start[lo]);
  const MultiDesc& d = m.probs[pidx];
  TB2DArgs a = d.a;
  const int cur = round & 1, nxt = cur ^ 1;
  a.ez_in = d.f[cur][0]; a.hx_in = d.f[cur][1]; a.hy_in = d.f[cur][2];
  a.ez_out = d.f[nxt][0]; a.hx_out = d.f[nxt][1]; a.hy_out = d.f[nxt][2];
  a.jd_in = d.jd[cur]; a.jd_out = d.jd[nxt];
  a.step0 = round * K;
  a.ksteps = min(K, a.T - a.step0);
  a.e_save = nullptr;
  // every problem of the batch is or is not given material ids (the host
  // sets them for all or none)
  mw_task<RW, K, POLES, MID, NW, false, true>(a, task, lane, w, x_e, x_h);
}

template <int RW, int K, bool MID, int NW = 4>
static hipError_t launch_round_mw_rk(const MultiArgs& m, int poles, int seg0, int seg1,
                                     long long first, long long count, hipStream_t s) {
  const dim3 blk(64 * NW, 1, 1), grd((unsigned)count, 1, 1);
  if (poles == 0) hipLaunchKernelGGL((round_mw_kernel<RW, K, 0, MID, NW>), grd, blk, 0, s, m, seg0, seg1, first, count);
  else if (poles == 1) hipLaunchKernelGGL((round_mw_kernel<RW, K, 1, MID, NW>), grd, blk, 0, s, m, seg0, seg1, first, count);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_round_mw(const MultiArgs& m, int poles, int R, int K, int NW, int seg0, int seg1,
                           long long first, long long count, bool mat_ids, hipStream_t s) {
  if (count < 1) return hipSuccess;
  if (count > 0x7fffffffLL) return hipErrorInvalidValue;
  if (NW == 8) {   // 64-row windows of eight 8-row wavefronts, depth 6 or 8
    if (R != 64) return hipErrorInvalidValue;
    if (K == 6)
      return mat_ids ? launch_round_mw_rk<8, 6, true, 8>(m, poles, seg0, seg1, first, count, s)
                     : launch_round_mw_rk<8, 6, false, 8>(m, poles, seg0, seg1, first, count, s);
    if (K == 8)
      return mat_ids ? launch_round_mw_rk<8, 8, true, 8>(m, poles, seg0, seg1, first, count, s)
                     : launch_round_mw_rk<8, 8, false, 8>(m, poles, seg0, seg1, first, count, s);
    return hipErrorInvalidValue;
  }
  if (NW != 4) return hipErrorInvalidValue;
#define CEEP_RMW(RW)                                                                         \
  switch (K) {                                                                               \
    case 4: return mat_ids ? launch_round_mw_rk<RW, 4, true>(m, poles, seg0, seg1, first, count, s)   \
                           : launch_round_mw_rk<RW, 4, false>(m, poles, seg0, seg1, first, count, s); \
    case 6: return mat_ids ? launch_round_mw_rk<RW, 6, true>(m, poles, seg0, seg1, first, count, s)   \
                           : launch_round_mw_rk<RW, 6, false>(m, poles, seg0, seg1, first, count, s); \
  }                                                                                          \
  return hipErrorInvalidValue;
  if (R == 32) { CEEP_RMW(8) }
  if (R == 48) { CEEP_RMW(12) }
#undef CEEP_RMW
  return hipErrorInvalidValue;
}

// Uneven batch in one persistent launch.  A workgroup is G groups of four
// wavefronts (G = 2..4, as many as the task's registers allow at 512 VGPRs
// per SIMD lane).  It takes whole items -- one problem's TX
// chunk, every round of it -- from an atomic counter over the item list
// (ordered largest first) and steps each alone: in a round the groups take
// the problem's tiles in turn, each tile being the four-wavefront task of
// step2d_mw_kernel (the same code, so every problem is bit-identical to its
// own run with that launch shape), and one workgroup barrier separates
// rounds (the ping-pong sets: a round reads what the last one wrote).  No
// workgroup ever waits for another, so the launch drains whatever the
// order in which workgroups become resident.  A group with no tile left in
// a round takes the task's barriers alone (two per step and TX plane), so
// every wavefront meets the same barriers.
template <int RW, int K, int POLES, bool MID, int G>
__global__ __launch_bounds__(256 * G) void persist_mw_kernel(PersistArgs m) {
  __shared__ float x_e[G][5][64], x_h[G][5][64];   // per group: the task's exchange rows
  __shared__ int item_s;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = wave >> 2, w = wave & 3;
  if (w == 0) {                                       // each group's edge rows (see mw_task)
    x_e[g][4][lane] = 0.f;
    x_h[g][0][lane] = 0.f;
  }
  for (;;) {
    if (threadIdx.x == 0) item_s = (int)atomicAdd(m.queue, 1u);
    __syncthreads();
    const int item = item_s;
    __syncthreads();                                  // read before the next take overwrites it
    if (item >= m.nitems) break;                      // uniform per workgroup
    const PersistItem it = m.items[item];
    const MultiDesc* d = m.probs + it.prob;
    const int tiles = d->a.tiles_x * d->a.tiles_y;
    const int nrounds = d->nrounds;
    for (int round = 0; round < nrounds; ++round) {
      for (int t0 = 0; t0 < tiles; t0 += G) {
        // the task's arguments are read afresh for each task (scalar loads)
        // rather than kept live across the loops
        TB2DArgs a = d->a;
        const int cur = round & 1, nxt = cur ^ 1;
        a.ez_in = d->f[cur][0]; a.hx_in = d->f[cur][1]; a.hy_in = d->f[cur][2];
        a.ez_out = d->f[nxt][0]; a.hx_out = d->f[nxt][1]; a.hy_out = d->f[nxt][2];
        a.jd_in = d->jd[cur]; a.jd_out = d->jd[nxt];
        a.step0 = round * K;
        a.ksteps = min(K, a.T - a.step0);
        a.e_save = nullptr;
        const int t = t0 + g;
        if (t < tiles) {
          mw_task<RW, K, POLES, MID>(a, it.chunk * tiles + t, lane, w, x_e[g], x_h[g]);
        } else {   // no tile left for this group: the task's barriers alone
          const int b_begin = it.chunk * a.bpw;
          const int nb = min(b_begin + a.bpw, a.B) - b_begin;
          for (int b = 0; b < nb; ++b)
            for (int st = 0; st < K && st < a.ksteps; ++st) {
              __syncthreads();
              __syncthreads();
            }
        }
      }
      __syncthreads();                                // this round's set is complete
    }
  }
}

// groups per workgroup: 32-row windows take 82 VGPRs (plain) and 131
// (Debye) per task, 48-row windows 110 and 182
constexpr int persist_groups(int RW, int poles) {
  return RW == 8 ? (poles ? 3 : 4) : (poles ? 2 : 3);
}

template <int RW, int K, bool MID>
static const void* persist_fn(int poles) {
  return poles == 0 ? (const void*)persist_mw_kernel<RW, K, 0, MID, persist_groups(RW, 0)>
                    : (const void*)persist_mw_kernel<RW, K, 1, MID, persist_groups(RW, 1)>;
}

static const void* persist_kernel(int poles, int R, int K, bool mat_ids) {
  if (poles != 0 && poles != 1) return nullptr;
#define CEEP_PMW(RW)                                                              \
  switch (K) {                                                                    \
    case 4: return mat_ids ? persist_fn<RW, 4, true>(poles) : persist_fn<RW, 4, false>(poles); \
    case 6: return mat_ids ? persist_fn<RW, 6, true>(poles) : persist_fn<RW, 6, false>(poles); \
  }                                                                               \
  return nullptr;
  if (R == 32) { CEEP_PMW(8) }
  if (R == 48) { CEEP_PMW(12) }
#undef CEEP_PMW
  return nullptr;
}

hipError_t launch_persist_mw(const PersistArgs& m, int poles, int R, int K, bool mat_ids,
                             int blocks, hipStream_t s) {
  const void* f = persist_kernel(poles, R, K, mat_ids);
  if (!f || blocks < 1 || m.nitems < 0) return hipErrorInvalidValue;
  void* args[] = {const_cast<PersistArgs*>(&m)};
  return hipLaunchKernel(f, dim3((unsigned)blocks), dim3(persist_threads(poles, R)), args, 0, s);
}

int persist_threads(int poles, int R) {
  return 256 * persist_groups(R == 32 ? 8 : 12, poles);
}

int persist_mw_blocks_per_cu(int poles, int R, int K, bool mat_ids) {
  const void* f = persist_kernel(poles, R, K, mat_ids);
  int n = 0;
  if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, persist_threads(poles, R), 0) !=
                hipSuccess)
    return 0;
  return n;
}

template <int RW, int K, int NW, bool FULL>
static hipError_t launch_mw_rkf(const TB2DArgs& a, int poles, hipStream_t s) {
  const dim3 blk(64 * NW, 1, 1), grd((unsigned)a.ntasks, 1, 1);
  const bool ids = a.mat_id != nullptr;
  if (poles == 0 && ids) hipLaunchKernelGGL((step2d_mw_kernel<RW, K, 0, true, NW, FULL>), grd, blk, 0, s, a);
  else if (poles == 0) hipLaunchKernelGGL((step2d_mw_kernel<RW, K, 0, false, NW, FULL>), grd, blk, 0, s, a);
  else if (poles == 1 && ids) hipLaunchKernelGGL((step2d_mw_kernel<RW, K, 1, true, NW, FULL>), grd, blk, 0, s, a);
  else if (poles == 1) hipLaunchKernelGGL((step2d_mw_kernel<RW, K, 1, false, NW, FULL>), grd, blk, 0, s, a);
  else if (poles == kPolesDiff)
    hipLaunchKernelGGL((step2d_mw_kernel<RW, K, kPolesDiff, false, NW, FULL>), grd, blk, 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// the launches of all K steps without the step loop's early exit; a run's
// last, shorter launch with it
template <int RW, int K, int NW>
static hipError_t launch_mw_rk(const TB2DArgs& a, int poles, hipStream_t s) {
  if (a.ksteps < 1 || a.ksteps > K) return hipErrorInvalidValue;
  return a.ksteps == K ? launch_mw_rkf<RW, K, NW, true>(a, poles, s)
                       : launch_mw_rkf<RW, K, NW, false>(a, poles, s);
}

template <int RW>
static hipError_t launch_mw_r(const TB2DArgs& a, int poles, int K, hipStream_t s) {
  switch (K) {
    case 4: return launch_mw_rk<RW, 4, 4>(a, poles, s);
    case 6: return launch_mw_rk<RW, 6, 4>(a, poles, s);
    case 8: return launch_mw_rk<RW, 8, 4>(a, poles, s);
  }
  return hipErrorInvalidValue;
}

bool step2d_mw_shape_ok(int R, int K, int NW) {
  if (NW == 4) return (R == 32 || R == 48 || R == 64) && (K == 4 || K == 6 || K == 8);
  if (NW == 8) return (R == 32 && K == 4) || (R == 48 && (K == 4 || K == 6)) ||
                      (R == 64 && (K == 4 || K == 6 || K == 8));
  return false;
}

hipError_t launch_step2d_mw(const TB2DArgs& a, int poles, int R, int K, int NW, hipStream_t s) {
  if (a.ntasks < 1) return hipSuccess;
  if (!step2d_mw_shape_ok(R, K, NW)) return hipErrorInvalidValue;
  if (NW == 8) {
    if (R == 32) return launch_mw_rk<4, 4, 8>(a, poles, s);
    if (R == 48) return K == 4 ? launch_mw_rk<6, 4, 8>(a, poles, s) : launch_mw_rk<6, 6, 8>(a, poles, s);
    switch (K) {
      case 4: return launch_mw_rk<8, 4, 8>(a, poles, s);
      case 6: return launch_mw_rk<8, 6, 8>(a, poles, s);
      case 8: return launch_mw_rk<8, 8, 8>(a, poles, s);
    }
    return hipErrorInvalidValue;
  }
  switch (R) {
    case 32: return launch_mw_r<8>(a, poles, K, s);
    case 48: return launch_mw_r<12>(a, poles, K, s);
    case 64: return launch_mw_r<16>(a, poles, K, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_step2d_tb(const TB2DArgs& a, int poles, int R, int K, hipStream_t s) {
  switch (R) {
    case 16: return launch_r<16>(a, poles, K, s);
    case 24: return launch_r<24>(a, poles, K, s);
    case 32: return launch_r<32>(a, poles, K, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ceep
