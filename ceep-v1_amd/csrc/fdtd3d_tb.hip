// Temporally blocked batched 3D Yee step for gfx950: K = 1 or 2 full steps
// per launch, each workgroup marching along x.
//
// Same update as h3d_kernel + e3d_kernel (fdtd_3d_batched.py:753-884,
// 949-958; ADE of fdtd_2d.py:163-178), same operand order and, like them,
// no floating-point contraction, so the state is bit-identical to the
// one-step kernels'.
//
// Mapping: a 1024-thread workgroup owns a 16 (y) x 64 (z) window of one TX:
// wavefront w holds row y0 + w, lane l holds z0 + l.  It walks the x-slabs
// in order with a K-level pipeline: in iteration i, level k (1..K) advances
// slab i - k + 1 from step s + k - 1 to s + k, so level k reads the slab
// level k - 1 finished in the same iteration (its +x neighbour) and keeps
// the H it produced one slab back (its -x neighbour).  z-neighbours come from
// DPP lane shifts, y-neighbours from the adjacent wavefronts through LDS
// (one barrier for the E rows of all levels, one per level for its H rows).
// Each level's window loses one row / lane of validity per side per step,
// so the (16 - 2K) x (64 - 2K) inner cells of level K are exact and are the
// only ones written.  The slab i + 2 loads are issued at the top of iteration
// i, a full iteration ahead of their use, and the barriers wait on LDS only.
//
// Cells: every state array is read from set `in_set` and written to the
// other set (ping-pong: neighbouring workgroups read the halo of a slab
// after this one has written it).  Slab Mx-1, row My-1 and column Mz-1 stay
// zero in every step (they hold walls only, and their H has zero curl), so
// they are neither stepped nor written.
#include "fdtd_common.h"
#include "fdtd3d_kernels.h"

namespace ceep {
namespace {

constexpr int kRows = 16;   // window rows: one wavefront each

template <int POLES>
struct Cell3 {
  float e[3], h[3];
  float jd[(POLES & 1) ? 3 : 1];
  float jl[(POLES & 2) ? 3 : 1];
  float jlp[(POLES & 2) ? 3 : 1];
};

struct Maps3 {
  float ca, cb, dk1, dk2, lk1, lk2, lk3, da, dbi;
  int ph;
  int id;   // material-id form: the cell's table row (the map values come from LDS at use)
};

// Float j of a material's two-float4 table row
__device__ __forceinline__ float tab_float(const float4& t0, const float4& t1, int j) {
  switch (j) {
    case 0: return t0.x;
    case 1: return t0.y;
    case 2: return t0.z;
    case 3: return t0.w;
    case 4: return t1.x;
    case 5: return t1.y;
    case 6: return t1.z;
    default: return t1.w;
  }
}

// The maps of table row m.id (the rest of m kept), read from the LDS copy of
// the material table; the record order is the packed maps' (map_pack)
template <int POLES>
__device__ __forceinline__ Maps3 resolve_maps(const float4* mtab, Maps3 m) {
  constexpr MapPack MPK = map_pack(POLES, false);
  const float4 t0 = mtab[2 * m.id], t1 = mtab[2 * m.id + 1];
  m.ca = t0.x;
  m.cb = t0.y;
  if (POLES & 1) {
    m.dk1 = tab_float(t0, t1, MPK.dk);
    m.dk2 = tab_float(t0, t1, MPK.dk + 1);
  }
  if (POLES & 2) {
    m.lk1 = tab_float(t0, t1, MPK.lk);
    m.lk2 = tab_float(t0, t1, MPK.lk + 1);
    m.lk3 = tab_float(t0, t1, MPK.lk + 2);
  }
  return m;
}

}  // namespace

template <int K, int POLES, bool MAG, bool MID>
__global__ __launch_bounds__(1024) void step3d_tb_kernel(TB3DArgs p) {
#pragma clang fp contract(off)
  static_assert(!(MID && MAG), "the material-id form has no mu_r / sigma_m maps");
  constexpr int R = kRows;
  constexpr int NJ = ((POLES & 1) ? 3 : 0) + ((POLES & 2) ? 6 : 0);
  constexpr int JL = (POLES & 1) ? 3 : 0;   // packed index of Lorentz J x
  __shared__ float xe[K][2][R][64];   // E (x, z) each level starts from: +y neighbours
  __shared__ float xh[K][2][R][64];   // H (z, x) after each level's H update: -y neighbours
  __shared__ float4 mtab[MID ? 2 * kMaxMat3 : 1];   // material table (MID)
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int L = xcd_block(p.order != 0);
  int ty, tz, bt;
  if (p.order == 2) {   // TX fastest: the TX of one tile share its maps in an L2
    bt = L % p.nb;
    const int t = L / p.nb;
    ty = t % p.tiles_y;
    tz = t / p.tiles_y;
  } else if (p.order == 3) {   // groups of tx_group TX fastest, then y tiles: an XCD
    const int g = p.tx_group;  // holds a few TX x a run of y tiles (maps and halos shared)
    const int r = L / g;
    ty = r % p.tiles_y;
    const int r2 = r / p.tiles_y;
    tz = r2 % p.tiles_z;
    bt = (r2 / p.tiles_z) * g + L % g;
    if (bt >= p.nb) return;   // padding of the last group (uniform per workgroup)
  } else {              // y tiles fastest: neighbouring tiles share halo rows
    ty = L % p.tiles_y;
    const int r = L / p.tiles_y;
    tz = r % p.tiles_z;
    bt = r / p.tiles_z;
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
  // lane byte offsets: fields (window loads, inner-cell stores), packed currents, packed maps
  const unsigned yzb = yz_in ? yz * 4u : kOffNone;
  const unsigned yzs = own ? yzb : kOffNone;
  const unsigned yzj = yz_in ? yz * (4u * NJ) : kOffNone;
  const unsigned yzjs = own ? yzj : kOffNone;
  constexpr MapPack MPK = map_pack(POLES, MAG);
  const unsigned yzm = yz_in ? yz * (4u * MPK.n) : kOffNone;
  const unsigned yzp = own ? yz * 2u : kOffNone;   // probe heads: inner cells only

  // resources (the host keeps (nbox + bstride) * 4 and bstride * 36 below 2^31)
  const long long tb = (long long)bt * p.bstride;
  const unsigned fbytes = (unsigned)((p.nbox + p.bstride) * 4);   // both sets of this TX
  const unsigned in_b = (unsigned)(p.in_set * p.nbox * 4);
  const unsigned out_b = (unsigned)((p.in_set ^ 1) * p.nbox * 4);
  const unsigned sx_b = (unsigned)(p.sx * 4);
  const unsigned sxj_b = sx_b * NJ;
  const unsigned sxm_b = sx_b * MPK.n;
  __amdgpu_buffer_rsrc_t re[3], rh[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    re[q] = plane_rsrc(p.e[q] + tb, fbytes);
    rh[q] = plane_rsrc(p.h[q] + tb, fbytes);
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
  const bool probes = p.probe_head != nullptr;
  const __amdgpu_buffer_rsrc_t rp = plane_rsrc(
      probes ? reinterpret_cast<const float*>(p.probe_head) : p.mp, (unsigned)(p.bstride * 2));
  const long long sx = p.sx;
  if (MID) {   // the table, once per workgroup (read after the barrier below)
    const float4* mt = reinterpret_cast<const float4*>(p.mat_table);
    for (int t = threadIdx.x; t < 2 * p.n_mat; t += 1024) mtab[t] = mt[t];
    lds_barrier();
  }

  const long long sc = p.src_cell[bt];
  const int src_a = (int)(sc / sx);
  const bool src_col = yz_in && (long long)yz == sc - (long long)src_a * sx;
  float wv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) wv[k] = p.wave[(long long)bt * p.wave_stride_b + p.step + k];

  // Slabs outside [0, Mx-1) load as zeros (slab Mx-1 and beyond hold zeros;
  // slab -1 is outside the box): their lanes take the out-of-range offset.
  // E of a slab is loaded two iterations ahead of its level-1 step (level 1
  // of the slab before needs it), H and J one iteration ahead.
  auto slab_ok = [&](int a) { return a >= 0 && a < p.Mx - 1; };
  auto load_e = [&](Cell3<POLES>& c, int a) {
    const unsigned v = slab_ok(a) ? yzb : kOffNone;
    const unsigned so = in_b + (unsigned)max(a, 0) * sx_b;
#pragma unroll
    for (int q = 0; q < 3; ++q) c.e[q] = rload(re[q], v, so);
  };
  auto load_hj = [&](Cell3<POLES>& c, int a) {
    const bool ok = slab_ok(a);
    const unsigned v = ok ? yzb : kOffNone, vj = ok ? yzj : kOffNone;
    const unsigned so = in_b + (unsigned)max(a, 0) * sx_b;
    const unsigned sj = (unsigned)max(a, 0) * sxj_b;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      c.h[q] = rload(rh[q], v, so);
      if (POLES & 1) c.jd[q] = rload(rj_in, vj + 4u * q, sj);
      if (POLES & 2) {
        c.jl[q] = rload(rj_in, vj + 4u * (JL + q), sj);
        c.jlp[q] = rload(rj_in, vj + 4u * (JL + 3 + q), sj);
      }
    }
  };
  auto load_maps = [&](Maps3& m, int a) {
    const bool ok = slab_ok(a);
    if (MID) {   // out-of-range slabs and cells read id 0, the all-zero row
      m.id = (int)__builtin_amdgcn_raw_buffer_load_b8(rid, (int)(ok ? yzi : kOffNone),
                                                      (int)((unsigned)max(a, 0) * (unsigned)sx), 0);
      m.ph = probes ? rload_short(rp, ok ? yzp : kOffNone, (unsigned)max(a, 0) * (unsigned)(p.sx * 2))
                    : -1;
      return;
    }
    const unsigned v = ok ? yzm : kOffNone;
    const unsigned so = (unsigned)max(a, 0) * sxm_b;
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
    // probe head: read (and masked to the inner cells) where it is used, so
    // this prefetch is not waited for here
    m.ph = probes ? rload_short(rp, ok ? yzp : kOffNone, (unsigned)max(a, 0) * (unsigned)(p.sx * 2))
                  : -1;
  };

  // Ring of four slab slots: slab x lives in ring[x % 4].  In iteration i
  // the slab loaded ahead is i + 2, level k works on slab i + 1 - k (the
  // slab of "level 0" is i + 1, already loaded).  Maps of slab x in mring[x % 4].
  Cell3<POLES> ring[4];
  Maps3 mring[4];
  float hpy[K + 1], hpz[K + 1];   // level k's H (y, z) one slab back: the -x neighbours
#pragma unroll
  for (int k = 0; k <= K; ++k) { hpy[k] = 0.f; hpz[k] = 0.f; }
  load_e(ring[3], -1);   // slab -1: zeros
  load_hj(ring[3], -1);
  load_maps(mring[3], -1);
  load_e(ring[0], 0);
  load_hj(ring[0], 0);
  load_maps(mring[0], 0);
  load_e(ring[1], 1);

  auto iteration = [&](auto phase, int i) {
    constexpr int U = decltype(phase)::value;   // i % 4
    load_e(ring[(U + 2) % 4], i + 2);
    load_hj(ring[(U + 1) % 4], i + 1);
    load_maps(mring[(U + 1) % 4], i + 1);
    // E (x, z) of every level's slab, for the +y neighbours
#pragma unroll
    for (int k = 1; k <= K; ++k) {
      xe[k - 1][0][w][lane] = ring[(U + 5 - k) % 4].e[0];
      xe[k - 1][1][w][lane] = ring[(U + 5 - k) % 4].e[2];
    }
    lds_barrier();
    float exy[K + 1], ezy[K + 1];
    {
      const int wn = (w + 1) & (R - 1);
#pragma unroll
      for (int k = 1; k <= K; ++k) {
        exy[k] = xe[k - 1][0][wn][lane];
        ezy[k] = xe[k - 1][1][wn][lane];
      }
    }
#pragma unroll
    for (int k = 1; k <= K; ++k) {
      const int a = i - k + 1;
      Cell3<POLES>& c = ring[(U + 5 - k) % 4];
      const Cell3<POLES>& cn = ring[(U + 6 - k) % 4];   // slab a + 1, at level k - 1
      const Maps3 m = MID ? resolve_maps<POLES>(mtab, mring[(U + 5 - k) % 4])
                          : mring[(U + 5 - k) % 4];
      const bool xa = a < p.Mx - 1;
      // H (h3d_kernel): forward differences
      {
        const float e0 = c.e[0], e1 = c.e[1], e2 = c.e[2];
        const float eyz = lane_next(e1), exz = lane_next(e0);
        const float ezx = cn.e[2], eyx = cn.e[1];
        const float da = m.da, cc = MAG ? m.dbi : p.ch;
        if (hx_ok) {
          const float t = (eyz - e1) - (ezy[k] - e2);
          c.h[0] = MAG ? da * c.h[0] + cc * t : c.h[0] + cc * t;
        }
        if (xa && hy_ok) {
          const float t = (ezx - e2) - (exz - e0);
          c.h[1] = MAG ? da * c.h[1] + cc * t : c.h[1] + cc * t;
        }
        if (xa && hz_ok) {
          const float t = (exy[k] - e0) - (eyx - e1);
          c.h[2] = MAG ? da * c.h[2] + cc * t : c.h[2] + cc * t;
        }
      }
      xh[k - 1][0][w][lane] = c.h[2];
      xh[k - 1][1][w][lane] = c.h[0];
      lds_barrier();
      const int ws = (w - 1) & (R - 1);
      const float hzy = xh[k - 1][0][ws][lane];
      const float hxy = xh[k - 1][1][ws][lane];
      // E (e3d_kernel) on the interior; -x neighbours kept from the last slab
      const float hx0 = c.h[0], hy0 = c.h[1], hz0 = c.h[2];
      const float hyz = lane_prev(hy0), hxz = lane_prev(hx0);
      const bool interior = e_yz && a >= 1 && a <= p.Mx - 2;
      if (interior) {
        float cu[3];
        cu[0] = (hz0 - hzy) - (hy0 - hyz);
        cu[1] = (hx0 - hxz) - (hz0 - hpz[k]);
        cu[2] = (hy0 - hpy[k]) - (hx0 - hxy);
        const float ca = m.ca, cb = m.cb;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const float e0 = c.e[q];
          if (POLES == 0) {
            c.e[q] = ca * e0 + cb * (cu[q] * p.idx);
          } else {
            float sgj = 0.f;
            if (POLES & 1) sgj += 0.5f * (1.f + m.dk1) * c.jd[q];
            if (POLES & 2) sgj += c.jl[q];
            const float en = ca * e0 + cb * (cu[q] * p.idx) - cb * sgj;
            const float de = en - e0;
            if (POLES & 1) c.jd[q] = m.dk1 * c.jd[q] + m.dk2 * de;
            if (POLES & 2) {
              const float jn = m.lk1 * c.jl[q] + m.lk3 * c.jlp[q] + m.lk2 * de;
              c.jlp[q] = c.jl[q];
              c.jl[q] = jn;
            }
            c.e[q] = en;
          }
        }
        if (src_col && a == src_a) c.e[2] += wv[k - 1];
        if (own && p.e_save) {
          float* es = p.e_save + (long long)(k - 1) * 3 * p.nbox + tb + (long long)a * sx + yz;
          es[0] = c.e[0];
          es[p.nbox] = c.e[1];
          es[2 * p.nbox] = c.e[2];
        }
      }
      if (probes && a >= 0 && a < p.Mx - 1)
        for (int q = own ? m.ph : -1; q >= 0; q = p.probe_next[q])
          p.traces[((long long)bt * p.P + q) * p.T + p.step + k - 1] = c.e[2];
      hpy[k] = hy0;
      hpz[k] = hz0;
    }
    // level K's slab is final
    const int a = i - K + 1;
    if (a >= 0 && a < p.Mx - 1) {
      const Cell3<POLES>& c = ring[(U + 5 - K) % 4];
      const unsigned so = out_b + (unsigned)a * sx_b;
      const unsigned sj = (unsigned)a * sxj_b;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        rstore(c.e[q], re[q], yzs, so);
        rstore(c.h[q], rh[q], yzs, so);
        if (POLES & 1) rstore(c.jd[q], rj_out, yzjs + 4u * q, sj);
        if (POLES & 2) {
          rstore(c.jl[q], rj_out, yzjs + 4u * (JL + q), sj);
          rstore(c.jlp[q], rj_out, yzjs + 4u * (JL + 3 + q), sj);
        }
      }
    }
  };

  // level K finishes slab Mx-2 at iteration Mx-3+K; the rounded-up tail
  // iterations only see zero slabs and write nothing
  const int n_iter = p.Mx - 2 + K;
  for (int i = 0; i < n_iter; i += 4) {
    iteration(Phase<0>{}, i);
    iteration(Phase<1>{}, i + 1);
    iteration(Phase<2>{}, i + 2);
    iteration(Phase<3>{}, i + 3);
  }
}

hipError_t launch_step3d_tb(const TB3DArgs& a, int K, int poles, hipStream_t s) {
  if (K != 1 && K != 2) return hipErrorInvalidValue;
  if (a.Mx < 3 || a.My < 3 || a.Mz < 3) return hipErrorInvalidValue;
  if (a.tiles_y != (a.My - 1 + (kRows - 2 * K) - 1) / (kRows - 2 * K) ||
      a.tiles_z != (a.Mz - 1 + (64 - 2 * K) - 1) / (64 - 2 * K))
    return hipErrorInvalidValue;
  if (a.nb < 1 || a.b0 < 0 || a.b0 + a.nb > a.B) return hipErrorInvalidValue;
  if (a.order == 3 && (a.tx_group < 1 || a.tx_group > a.nb)) return hipErrorInvalidValue;
  const long long nblk = (long long)a.tiles_y * a.tiles_z *
                         (a.order == 3 ? (a.nb + a.tx_group - 1) / a.tx_group * a.tx_group : a.nb);
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grd((unsigned)nblk, 1, 1), blk(1024, 1, 1);
  if (!a.mp || (poles != 0 && !a.jp)) return hipErrorInvalidValue;
  const bool mag = a.magnetic != 0;
  const bool mid = a.mat_id != nullptr;
  if (mid && (mag || !a.mat_table || a.n_mat < 1 || a.n_mat > kMaxMat3)) return hipErrorInvalidValue;
#define CEEP_TB3(KK, PP)                                                                     \
  do {                                                                                       \
    if (mag) hipLaunchKernelGGL((step3d_tb_kernel<KK, PP, true, false>), grd, blk, 0, s, a); \
    else if (mid) hipLaunchKernelGGL((step3d_tb_kernel<KK, PP, false, true>), grd, blk, 0, s, a); \
    else hipLaunchKernelGGL((step3d_tb_kernel<KK, PP, false, false>), grd, blk, 0, s, a);    \
  } while (0)
#define CEEP_TB3K(PP)            \
  do {                           \
    if (K == 2) CEEP_TB3(2, PP); \
    else CEEP_TB3(1, PP);        \
  } while (0)
  switch (poles) {
    case 0: CEEP_TB3K(0); break;
    case 1: CEEP_TB3K(1); break;
    case 2: CEEP_TB3K(2); break;
    case 3: CEEP_TB3K(3); break;
    default: return hipErrorInvalidValue;
  }
#undef CEEP_TB3K
#undef CEEP_TB3
  return hipGetLastError();
}

}  // namespace ceep
