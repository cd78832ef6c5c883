// Temporally blocked time-reversed 3D step for gfx950: K = 1 or 2 reversed
// steps per launch in place of K x (A, B, C) of fdtd3d_adjoint.hip, with the
// same operations in the same order (contraction off in both).
//
// Mapping as the blocked forward (fdtd3d_tb.hip): a 1024-thread workgroup
// owns a 16 (y) x 64 (z) window of one TX and walks the x-slabs in order.  In
// iteration i, level k (k = 1..K, reversed step S - k + 1) runs A on slab
// i + 2 - k (pointwise), then B and C on slab i + 1 - k: B reads q at +1 in x
// from the slab that level's A just finished, C reads r = dbi hbar at -1 in x
// kept from the slab before; y-neighbours pass through LDS (one barrier for
// q, one for r per level), z-neighbours through DPP lane shifts.  Level k + 1
// runs A on the slab level k has just finished with C.  Each level loses one
// cell of validity per side, so the inner (16 - 2K) x (64 - 2K) cells are
// written: the adjoint currents right after the last level's A, lam and hbar
// after its C.  The gradient sums of a slab are formed once its last A is
// done, from the et of every level and the K + 1 E-store slots.  Slab Mx-1,
// row My-1 and column Mz-1 hold zeros in every step.
//
// With K = 2, level 1's adjoint currents and et wait in LDS (one slot per
// thread, so no barrier) from its A on a slab to level 2's A on that slab an
// iteration later, which keeps the Debye + Lorentz case within 128 VGPRs.
// The E-store and gradient loads of a slab are issued before the next
// slab's prefetch: vector memory counters complete in order, so waiting for
// loads issued after the prefetch would wait for the prefetch as well.  The
// maps are the forward's packed records (map_pack).
//
// State: lam and hbar per component in two sets (one allocation of
// 2 * nbox floats each), the adjoint currents packed per cell in two sets
// ([set][B][cell][nj]: Debye x, y, z; Lorentz x, y, z; Lorentz one reversed
// step back x, y, z); a launch reads set in_set and writes the other.  The
// gradient sums go to one plane per TX and coefficient (g[2][B][cell],
// read-modify-write by the one workgroup that owns the cell).
#include "fdtd_common.h"
#include "fdtd3d_kernels.h"

namespace ceep {
namespace {

constexpr int kRowsA = 16;

template <int POLES, int K>
struct ASlab {
  float lam[3], hb[3], q[3];
  float ad[(POLES & 1) ? 3 : 1];
  float al[(POLES & 2) ? 3 : 1];
  float alp[(POLES & 2) ? 3 : 1];
  float et[K][3];   // the adjoint of each level's new E, for the gradient sums
};

template <int K>
struct EGSlab {
  float e[K + 1][3];   // E-store slots 0..K
  float gca, gcb;
};

struct AMaps {
  float ca, cb, dk1, dk2, lk1, lk2, lk3, da, dbi;
  int ph;
  int id;   // material-id form: the cell's table row (the map values come from LDS at use)
};

// The maps of table row m.id (the rest of m kept) from the LDS copy of the
// material table, two float4 per row in the packed maps' order (map_pack)
template <int POLES>
__device__ __forceinline__ AMaps resolve_amaps(const float4* mtab, AMaps m) {
  constexpr MapPack MPK = map_pack(POLES, false);
  const float4 t0 = mtab[2 * m.id], t1 = mtab[2 * m.id + 1];
  const float v[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
  m.ca = v[0];
  m.cb = v[1];
  if (POLES & 1) {
    m.dk1 = v[MPK.dk];
    m.dk2 = v[MPK.dk + 1];
  }
  if (POLES & 2) {
    m.lk1 = v[MPK.lk];
    m.lk2 = v[MPK.lk + 1];
    m.lk3 = v[MPK.lk + 2];
  }
  return m;
}

}  // namespace

template <int K, int POLES, bool MAG, bool MID>
__global__ __launch_bounds__(1024) void adj3d_tb_kernel(AdjTB3DArgs p) {
#pragma clang fp contract(off)
  static_assert(!(MID && MAG), "the material-id form has no mu_r / sigma_m maps");
  constexpr int R = kRowsA;
  constexpr int NJ = ((POLES & 1) ? 3 : 0) + ((POLES & 2) ? 6 : 0);
  constexpr int JL = (POLES & 1) ? 3 : 0;
  __shared__ float xq[K][2][R][64];   // q (x, z): B's +y neighbours
  __shared__ float xr[K][2][R][64];   // r (z, x): C's -y neighbours
  // K = 2: the adjoint currents and et of level 1 wait here, one slot per
  // thread, from level 1's A on a slab to level 2's A on it an iteration
  // later (two buffers by slab parity; each thread reads only its own slots)
  constexpr int NP = K == 2 ? NJ + 3 : 1;
  __shared__ float park[K == 2 ? 2 : 1][NP][1024];
  __shared__ float4 mtab[MID ? 2 * kMaxMat3 : 1];   // material table (MID)
  const int tid = threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int L = xcd_block(p.order != 0);
  int ty, tz, bt;
  if (p.order == 2) {   // TX fastest: the TX of one tile share its maps in an L2
    bt = L % p.nb;
    const int t = L / p.nb;
    ty = t % p.tiles_y;
    tz = t / p.tiles_y;
  } else {              // y tiles fastest
    ty = L % p.tiles_y;
    const int r0 = L / p.tiles_y;
    tz = r0 % p.tiles_z;
    bt = r0 / p.tiles_z;
  }
  bt += p.b0;
  const int y = ty * (R - 2 * K) - K + w;
  const int z = tz * (64 - 2 * K) - K + lane;
  const bool yz_in = y >= 0 && y < p.My && z >= 0 && z < p.Mz;
  const bool own = w >= K && w < R - K && lane >= K && lane < 64 - K && yz_in;
  const bool e_yz = y >= 1 && y <= p.My - 2 && z >= 1 && z <= p.Mz - 2;
  const bool hx_ok = y < p.My - 1 && z < p.Mz - 1;
  const bool hy_ok = z < p.Mz - 1;
  const bool hz_ok = y < p.My - 1;
  const unsigned yz = yz_in ? (unsigned)(y * p.pitch + z) : 0u;
  const unsigned yzb = yz_in ? yz * 4u : kOffNone;    // window loads
  const unsigned yzs = own ? yzb : kOffNone;          // inner-cell stores, E store, gradient
  const unsigned yzj = yz_in ? yz * (4u * NJ) : kOffNone;
  const unsigned yzjs = own ? yzj : kOffNone;
  constexpr MapPack MPK = map_pack(POLES, MAG);
  const unsigned yzm = yz_in ? yz * (4u * MPK.n) : kOffNone;
  // probe heads of every window cell: with K = 2 level 1's halo cells feed
  // level 2, so their lam needs its seeds too
  const unsigned yzp = yz_in ? yz * 2u : kOffNone;

  // resources (the host keeps (3 nbox + bstride) * 4 and bstride * 36 below 2^31)
  const long long tb = (long long)bt * p.bstride;
  const unsigned fbytes = (unsigned)((p.nbox + p.bstride) * 4);
  const unsigned in_b = (unsigned)(p.in_set * p.nbox * 4);
  const unsigned out_b = (unsigned)((p.in_set ^ 1) * p.nbox * 4);
  const unsigned nbox_b = (unsigned)(p.nbox * 4);
  const unsigned sx_b = (unsigned)(p.sx * 4);
  const unsigned sxj_b = sx_b * NJ;
  const unsigned sxm_b = sx_b * MPK.n;
  __amdgpu_buffer_rsrc_t rl[3], rh[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    rl[c] = plane_rsrc(p.lam[c] + tb, fbytes);
    rh[c] = plane_rsrc(p.hbar[c] + tb, fbytes);
  }
  const unsigned jbytes = (unsigned)(p.bstride * 4 * NJ);
  const __amdgpu_buffer_rsrc_t rj_in =
      plane_rsrc(NJ ? p.jp + ((long long)p.in_set * p.B + bt) * p.bstride * NJ : p.mp, jbytes);
  const __amdgpu_buffer_rsrc_t rj_out =
      plane_rsrc(NJ ? p.jp + ((long long)(p.in_set ^ 1) * p.B + bt) * p.bstride * NJ : p.mp,
                 jbytes);
  const __amdgpu_buffer_rsrc_t rm = plane_rsrc(p.mp, (unsigned)(p.bstride * 4 * MPK.n));
  // MID: one byte per cell; lane offset yz, slab offset a * sx
  const __amdgpu_buffer_rsrc_t rid = plane_rsrc(
      MID ? reinterpret_cast<const float*>(p.mat_id) : p.mp, (unsigned)p.bstride);
  const unsigned yzi = yz_in ? yz : kOffNone;
  // E-store slots of steps S - K + 1 .. S + 1, one resource each (component
  // c at + c * nbox; the host keeps (2 nbox + bstride) * 4 below 2^31)
  __amdgpu_buffer_rsrc_t re[K + 1];
#pragma unroll
  for (int j = 0; j <= K; ++j)
    re[j] = plane_rsrc(p.e_lo + (long long)j * 3 * p.nbox + tb,
                       (unsigned)((2 * p.nbox + p.bstride) * 4));
  // both gradient planes of the TX: g_cbi = g_ca + B * gstride (kept below 2^31 bytes)
  const __amdgpu_buffer_rsrc_t rg = plane_rsrc(
      p.g + (long long)bt * p.gstride, (unsigned)((p.B * p.gstride + p.gstride) * 4));
  const unsigned gcb_b = (unsigned)(p.B * p.gstride * 4);
  const bool probes = p.probe_head != nullptr;
  const __amdgpu_buffer_rsrc_t rp = plane_rsrc(
      p.probe_head ? reinterpret_cast<const float*>(p.probe_head) : p.mp, (unsigned)(p.bstride * 2));
  const long long sx = p.sx;
  if (MID) {   // the table, once per workgroup (read after the barrier below)
    const float4* mt = reinterpret_cast<const float4*>(p.mat_table);
    for (int t = tid; t < 2 * p.n_mat; t += 1024) mtab[t] = mt[t];
    lds_barrier();
  }

  const long long sc = p.src_cell[bt];
  const int src_a = (int)(sc / sx);
  const bool src_col = yz_in && (long long)yz == sc - (long long)src_a * sx;
  float inj_w[K];
#pragma unroll
  for (int k = 0; k < K; ++k) inj_w[k] = p.wave[(long long)bt * p.wave_stride_b + p.step - k];

  auto slab_ok = [&](int a) __attribute__((always_inline)) { return a >= 0 && a < p.Mx - 1; };
  auto load_slab = [&](ASlab<POLES, K>& c, int a) __attribute__((always_inline)) {
    const bool ok = slab_ok(a);
    const unsigned v = ok ? yzb : kOffNone, vj = ok ? yzj : kOffNone;
    const unsigned so = in_b + (unsigned)max(a, 0) * sx_b;
    const unsigned sj = (unsigned)max(a, 0) * sxj_b;
#pragma unroll
    for (int c3 = 0; c3 < 3; ++c3) {
      c.lam[c3] = rload(rl[c3], v, so);
      c.hb[c3] = rload(rh[c3], v, so);
      if (POLES & 1) c.ad[c3] = rload(rj_in, vj + 4u * c3, sj);
      if (POLES & 2) {
        c.al[c3] = rload(rj_in, vj + 4u * (JL + c3), sj);
        c.alp[c3] = rload(rj_in, vj + 4u * (JL + 3 + c3), sj);
      }
    }
  };
  auto load_maps = [&](AMaps& m, int a) __attribute__((always_inline)) {
    const bool ok = slab_ok(a);
    if (MID) {   // out-of-range slabs and cells read id 0, the all-zero row
      m.id = (int)__builtin_amdgcn_raw_buffer_load_b8(rid, (int)(ok ? yzi : kOffNone),
                                                      (int)((unsigned)max(a, 0) * (unsigned)sx), 0);
      m.dk1 = m.dk2 = m.lk1 = m.lk2 = m.lk3 = 0.f;
      m.da = 1.f;
      m.dbi = p.ch;
      m.ph = probes ? rload_short(rp, ok ? yzp : kOffNone, (unsigned)max(a, 0) * (unsigned)(sx * 2))
                    : -1;
      return;
    }
    const unsigned v = ok ? yzm : kOffNone;
    const unsigned so = (unsigned)max(a, 0) * sxm_b;
    m.dk1 = m.dk2 = m.lk1 = m.lk2 = m.lk3 = 0.f;
    m.da = 1.f;
    m.dbi = p.ch;
    m.ca = rload(rm, v, so);
    m.cb = rload(rm, v + 4u, so);
    if (POLES & 1) {
      m.dk1 = rload(rm, v + 4u * MPK.dk, so);
      m.dk2 = rload(rm, v + 4u * (MPK.dk + 1), so);
    }
    if (POLES & 2) {
      m.lk1 = rload(rm, v + 4u * MPK.lk, so);
      m.lk2 = rload(rm, v + 4u * (MPK.lk + 1), so);
      m.lk3 = rload(rm, v + 4u * (MPK.lk + 2), so);
    }
    if (MAG) {
      m.da = rload(rm, v + 4u * MPK.mag, so);
      m.dbi = rload(rm, v + 4u * (MPK.mag + 1), so);
    }
    m.ph = probes ? rload_short(rp, ok ? yzp : kOffNone, (unsigned)max(a, 0) * (unsigned)(sx * 2))
                  : -1;
  };

  // A (adj3d_a_kernel) of level k on slab a: pointwise
  auto step_a = [&](ASlab<POLES, K>& c, const AMaps& m, int a, int k) __attribute__((always_inline)) {
    const bool interior = e_yz && a >= 1 && a <= p.Mx - 2;
    const int pb = a & 1;
    if (K == 2 && k == 1) {
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        if (POLES & 1) c.ad[c3] = park[pb][c3][tid];
        if (POLES & 2) {
          c.al[c3] = park[pb][JL + c3][tid];
          c.alp[c3] = park[pb][JL + 3 + c3][tid];
        }
      }
    }
    if (interior) {
      const float ca = m.ca, cb = m.cb, cbi = cb * p.idx;
      float gam = 0.f;
      if (POLES & 1) gam = 0.5f * (1.f + m.dk1);
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        float G = 0.f;
        if (POLES & 1) G += m.dk2 * c.ad[c3];
        if (POLES & 2) G += m.lk2 * c.al[c3];
        const float et = c.lam[c3] + G;
        c.et[k][c3] = et;
        c.lam[c3] = ca * et - G;
        c.q[c3] = cbi * et;
        if (POLES & 1) c.ad[c3] = m.dk1 * c.ad[c3] - cb * gam * et;
        if (POLES & 2) {
          const float an = m.lk1 * c.al[c3] + m.lk3 * c.alp[c3] - cb * et;
          c.alp[c3] = c.al[c3];
          c.al[c3] = an;
        }
      }
    } else {   // E never changes off the interior: no adjoint flows through it
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) { c.lam[c3] = 0.f; c.q[c3] = 0.f; c.et[k][c3] = 0.f; }
    }
    if (K == 2 && k == 0) {
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        if (POLES & 1) park[pb][c3][tid] = c.ad[c3];
        if (POLES & 2) {
          park[pb][JL + c3][tid] = c.al[c3];
          park[pb][JL + 3 + c3][tid] = c.alp[c3];
        }
        park[pb][NJ + c3][tid] = c.et[0][c3];
      }
    }
  };
  // after the last level's A on slab a: store its adjoint currents, and add
  // the gradient sums of every level (level k is reversed step S - k, which
  // reads E-store slots S - k and S - k + 1; slot j of the launch is step
  // S - K + 1 + j)
  // the E-store slots and gradient sums of slab a, issued ahead of use
  // (vector memory counters complete in order: loads issued after the next
  // slab's prefetch would make their wait wait for the prefetch too)
  auto load_eg = [&](EGSlab<K>& g, int a) __attribute__((always_inline)) {
    const unsigned v = slab_ok(a) ? yzs : kOffNone;
    const unsigned sa = (unsigned)max(a, 0) * sx_b;
    g.gca = rload(rg, v, sa);
    g.gcb = rload(rg, v, sa + gcb_b);
#pragma unroll
    for (int j = 0; j <= K; ++j)
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) g.e[j][c3] = rload(re[j], v, sa + c3 * nbox_b);
  };
  auto finish_a = [&](const ASlab<POLES, K>& c, const AMaps& m, int a,
                      const EGSlab<K>& g) __attribute__((always_inline)) {
    if (!slab_ok(a)) return;
    const unsigned sa = (unsigned)a * sx_b;
    const unsigned sj = (unsigned)a * sxj_b;
#pragma unroll
    for (int c3 = 0; c3 < 3; ++c3) {
      if (POLES & 1) rstore(c.ad[c3], rj_out, yzjs + 4u * c3, sj);
      if (POLES & 2) {
        rstore(c.al[c3], rj_out, yzjs + 4u * (JL + c3), sj);
        rstore(c.alp[c3], rj_out, yzjs + 4u * (JL + 3 + c3), sj);
      }
    }
    float gca = g.gca, gcb = g.gcb;
    if (e_yz && a >= 1 && a <= p.Mx - 2) {
      const float(&e)[K + 1][3] = g.e;
      const bool at_src = src_col && a == src_a;
      const int pb = a & 1;
#pragma unroll
      for (int k = 0; k < K; ++k) {   // level k + 1: step S - k, slots K - 1 - k and K - k
        float et[3];
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) et[c3] = (K == 2 && k == 0) ? park[pb][NJ + c3][tid] : c.et[k][c3];
        const float* es = e[K - 1 - k];
        const float* es1 = e[K - k];
        const float inj = at_src ? inj_w[k] : 0.f;
        float sa_ = 0.f, sb_ = 0.f;
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          sa_ += et[c3] * es[c3];
          sb_ += et[c3] * (es1[c3] - (c3 == 2 ? inj : 0.f) - m.ca * es[c3]);
        }
        gca += sa_;
        gcb += sb_;
      }
    }
    rstore(gca, rg, yzs, sa);
    rstore(gcb, rg, yzs, sa + gcb_b);
  };

  ASlab<POLES, K> ring[4];
  AMaps mring[4];
  float rpy[K], rpz[K];   // r (y, z) of each level's previous slab: C's -x neighbours
#pragma unroll
  for (int k = 0; k < K; ++k) { rpy[k] = 0.f; rpz[k] = 0.f; }
  load_slab(ring[0], 0);
  load_maps(mring[0], 0);
  load_slab(ring[1], 1);
  load_maps(mring[1], 1);
  load_slab(ring[3], -1);   // slab -1: zeros (K = 2: level 2's first B and C
  load_maps(mring[3], -1);  // run on it; their results are never stored)
#pragma unroll
  for (int c3 = 0; c3 < 3; ++c3) ring[3].q[c3] = 0.f;
  // slab x's E-store slots and gradient sums live in eg[x & 1]: K = 2 loads
  // them at the start of the iteration that finishes x, K = 1 one iteration
  // earlier with the slab prefetch
  EGSlab<K> eg[2];
  if (K == 1) {
    load_eg(eg[0], 0);
    load_eg(eg[1], 1);
  }
  auto maps_of = [&](const AMaps& m) __attribute__((always_inline)) {
    return MID ? resolve_amaps<POLES>(mtab, m) : m;
  };
  {
    const AMaps m0 = maps_of(mring[0]);
    step_a(ring[0], m0, 0, 0);
    if (K == 1) finish_a(ring[0], m0, 0, eg[0]);
  }

  auto iteration = [&](auto phase, int i) __attribute__((always_inline)) {
    constexpr int U = decltype(phase)::value;   // i % 4: slab x lives in ring[x % 4]
    if (K == 2) load_eg(eg[U & 1], i);
#pragma unroll
    for (int k = 0; k < K; ++k) {       // level k + 1: reversed step S - k
      const int xa = i + 1 - k;         // A's slab
      const int xb = i - k;             // B's and C's slab
      ASlab<POLES, K>& sA = ring[(U + 5 - k) % 4];
      ASlab<POLES, K>& sB = ring[(U + 4 - k) % 4];
      const AMaps mA = maps_of(mring[(U + 5 - k) % 4]);
      step_a(sA, mA, xa, k);
      if (k == K - 1) {
        finish_a(sA, mA, xa, eg[(U + 1 - k) & 1]);
        // the next slab's prefetch, issued after every load the rest of
        // this iteration waits for
        load_slab(ring[(U + 2) % 4], i + 2);
        load_maps(mring[(U + 2) % 4], i + 2);
        if (K == 1) load_eg(eg[U & 1], i + 2);
      }
      // B: q at +1 in z (DPP), y (LDS), x (slab xb + 1 = A's)
      xq[k][0][w][lane] = sB.q[0];
      xq[k][1][w][lane] = sB.q[2];
      lds_barrier();
      {
        const int wn = (w + 1) & (R - 1);
        const float qxy = xq[k][0][wn][lane], qzy = xq[k][1][wn][lane];
        const float qx0 = sB.q[0], qy0 = sB.q[1], qz0 = sB.q[2];
        const float qyz = lane_next(qy0), qxz = lane_next(qx0);
        const float qzx = sA.q[2], qyx = sA.q[1];
        const bool xok = xb < p.Mx - 1;
        const float da = mring[(U + 4 - k) % 4].da;
        if (hx_ok) {
          const float t = (qy0 - qyz) - (qz0 - qzy);
          sB.hb[0] = MAG ? da * sB.hb[0] + t : sB.hb[0] + t;
        }
        if (xok && hy_ok) {
          const float t = (qz0 - qzx) - (qx0 - qxz);
          sB.hb[1] = MAG ? da * sB.hb[1] + t : sB.hb[1] + t;
        }
        if (xok && hz_ok) {
          const float t = (qx0 - qxy) - (qy0 - qyx);
          sB.hb[2] = MAG ? da * sB.hb[2] + t : sB.hb[2] + t;
        }
      }
      // C: r = dbi hbar at -1 in z (DPP), y (LDS), x (the level's previous slab)
      const AMaps& mB = mring[(U + 4 - k) % 4];
      const float c0 = MAG ? mB.dbi : p.ch;
      const float rx0 = c0 * sB.hb[0], ry0 = c0 * sB.hb[1], rz0 = c0 * sB.hb[2];
      xr[k][0][w][lane] = rz0;
      xr[k][1][w][lane] = rx0;
      lds_barrier();
      {
        const int ws = (w - 1) & (R - 1);
        const float rz_my = xr[k][0][ws][lane], rx_my = xr[k][1][ws][lane];
        const float ry_mz = lane_prev(ry0), rx_mz = lane_prev(rx0);
        if (e_yz && xb >= 1 && xb <= p.Mx - 2) {
          sB.lam[0] += (ry0 - ry_mz) - (rz0 - rz_my);
          sB.lam[1] += (rz0 - rpz[k]) - (rx0 - rx_mz);
          float lz = sB.lam[2] + ((rx0 - rx_my) - (ry0 - rpy[k]));
          const int st = p.step - k;    // this level's reversed step: seed dL/dtrace(st - 1)
          if (probes && st >= 1)       // (an interior cell: its probe head was loaded)
            for (int q = mB.ph; q >= 0; q = p.probe_next[q])
              lz += p.dtr[((long long)bt * p.P + q) * p.T + st - 1];
          sB.lam[2] = lz;
        }
      }
      rpy[k] = ry0;
      rpz[k] = rz0;
      if (k == K - 1 && slab_ok(xb)) {   // lam and hbar of slab xb are final
        const unsigned so = out_b + (unsigned)xb * sx_b;
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          rstore(sB.lam[c3], rl[c3], yzs, so);
          rstore(sB.hb[c3], rh[c3], yzs, so);
        }
      }
    }
  };

  // the last level's C finishes slab Mx-2 in iteration Mx-2+K-1; the
  // rounded-up tail only meets zero slabs
  const int n_iter = p.Mx - 2 + K;
  for (int i = 0; i < n_iter; i += 4) {
    iteration(Phase<0>{}, i);
    iteration(Phase<1>{}, i + 1);
    iteration(Phase<2>{}, i + 2);
    iteration(Phase<3>{}, i + 3);
  }
}

hipError_t launch_adj3d_tb(const AdjTB3DArgs& a, int K, int poles, hipStream_t s) {
  if (K != 1 && K != 2) return hipErrorInvalidValue;
  if (a.Mx < 3 || a.My < 3 || a.Mz < 3) return hipErrorInvalidValue;
  if (a.tiles_y != (a.My - 1 + (kRowsA - 2 * K) - 1) / (kRowsA - 2 * K) ||
      a.tiles_z != (a.Mz - 1 + (64 - 2 * K) - 1) / (64 - 2 * K))
    return hipErrorInvalidValue;
  if (!a.mp || (poles != 0 && !a.jp) || !a.e_lo || !a.g || a.step < K - 1)
    return hipErrorInvalidValue;
  if (a.nb < 1 || a.b0 < 0 || a.b0 + a.nb > a.B) return hipErrorInvalidValue;
  const long long nblk = (long long)a.tiles_y * a.tiles_z * a.nb;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grd((unsigned)nblk, 1, 1), blk(1024, 1, 1);
  const bool mag = a.magnetic != 0;
  const bool mid = a.mat_id != nullptr;
  if (mid && (mag || !a.mat_table || a.n_mat < 1 || a.n_mat > kMaxMat3)) return hipErrorInvalidValue;
#define CEEP_ATB(KK, PP)                                                                    \
  do {                                                                                      \
    if (mag) hipLaunchKernelGGL((adj3d_tb_kernel<KK, PP, true, false>), grd, blk, 0, s, a); \
    else if (mid) hipLaunchKernelGGL((adj3d_tb_kernel<KK, PP, false, true>), grd, blk, 0, s, a); \
    else hipLaunchKernelGGL((adj3d_tb_kernel<KK, PP, false, false>), grd, blk, 0, s, a);    \
  } while (0)
#define CEEP_ATBK(PP)            \
  do {                           \
    if (K == 2) CEEP_ATB(2, PP); \
    else CEEP_ATB(1, PP);        \
  } while (0)
  switch (poles) {
    case 0: CEEP_ATBK(0); break;
    case 1: CEEP_ATBK(1); break;
    case 2: CEEP_ATBK(2); break;
    case 3: CEEP_ATBK(3); break;
    default: return hipErrorInvalidValue;
  }
#undef CEEP_ATBK
#undef CEEP_ATB
  return hipGetLastError();
}

}  // namespace ceep
