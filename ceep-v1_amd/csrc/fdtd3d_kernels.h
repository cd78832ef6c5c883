// Launch interface of the 3D step kernels (device pointers, FP32).
#pragma once
#include <hip/hip_runtime.h>

#include "fdtd_common.h"

namespace ceep {

struct Step3DArgs {
  int Mx, My, Mz, pitch, B;       // cavity box, z pitch (floats), batch
  long long sx;                   // floats per x-slab (My * pitch)
  long long bstride;              // floats per batch volume (Mx * sx)
  float* ex; float* ey; float* ez;
  float* hx; float* hy; float* hz;
  float* jd[3];                   // Debye J per E component (in place)
  float* jl_cur[3];               // Lorentz J per component
  float* jl_prev[3];              // Lorentz J one step back; receives J_next
  const float* ca;
  const float* cb;
  const float* da;                // null -> da = 1, db * inv = ch (uniform)
  const float* dbi;               // db / dx per cell (with da)
  const float* d_k1;
  const float* d_k2dt;
  const float* l_k1;
  const float* l_k2dt;
  const float* l_k3;
  float ch;                       // dt / (mu0 dx) when da == null
  float idx;                      // 1 / dx
  const long long* src_cell;      // [B] box-linear index of each TX
  const float* wave;
  long long wave_stride_b;
  int step, T, P;
  const short* probe_head;
  const short* probe_next;
  float* traces;                  // [B][P][T]
  float* e_save;                  // if set: E after this step, [3][B][vol] (adjoint store)
  long long e_cstride;            // floats per component block of e_save (B * bstride)
  int xcd_order;                  // H/E kernels: block order (0 dispatch, 1/2 per XCD)
};

// Physical-CPML mode of the 3D engine: the full grid, fields in place, the
// reference's single-run 3D update (fdtd_3d.py:97-171) with its convolutional
// PML (absorbing.py:567-943).  psi arrays use the field layout; only the
// cells of their slabs are read or written.
struct Cpml3DArgs {
  int Mx, My, Mz, pitch, B, t;    // full grid, CPML thickness t
  long long sx, bstride;
  float* ex; float* ey; float* ez;
  float* hx; float* hy; float* hz;
  float* jd[3]; float* jl_cur[3]; float* jl_prev[3];
  // psi_<field><axis>: the recursion of that field's correction along that axis
  float* psi_hxy; float* psi_hxz; float* psi_hyz; float* psi_hyx; float* psi_hzx; float* psi_hzy;
  float* psi_exy; float* psi_exz; float* psi_eyz; float* psi_eyx; float* psi_ezx; float* psi_ezy;
  const float* ca; const float* cb;
  const float* da;                // null -> da = 1, dbi = ch
  const float* dbi;               // db / dx per cell
  const float* d_k1; const float* d_k2dt;
  const float* l_k1; const float* l_k2dt; const float* l_k3;
  float ch, idx;                  // (dt / mu0) / dx, 1 / dx
  float dtmu, dteps;              // dt / mu0, dt / eps0 (the psi corrections)
  float bh[kMaxCpml], chh[kMaxCpml], be[kMaxCpml], ce[kMaxCpml];
  const long long* src_cell;
  const float* wave;
  long long wave_stride_b;
  int step, T, P;
  const short* probe_head;
  const short* probe_next;
  float* traces;
};

// One step: H + H psi, then E (+ADE) + Ez source + E psi + Ez probes.
hipError_t launch_cpml3d(const Cpml3DArgs& a, int poles, hipStream_t s);

// One adjoint (time-reversed) step of the 3D update with the Debye/Lorentz
// ADE, as three kernels over the adjoint state (lam = adjoint of E, hbar =
// adjoint of H, ad/al = adjoints of J_Debye, J_Lorentz; the adjoint of
// J_Lorentz_prev is k3L times al one reversed step back, so al is kept in two
// buffers that swap roles every step instead of storing that product):
//   A (pointwise): G = k2D ad + k2L al;  Et = lam + G;  lam <- ca Et - G;
//       q = cb/dx Et;  ad <- k1D ad - cb gammaD Et;
//       al_prev <- k1L al + k3L al_prev - cb Et  (the new al);
//       g_ca += Et e^s;  g_cbi += Et (e^{s+1} - source - ca e^s)
//   B: hbar <- da hbar + curl^T q        (q taken at the +1 neighbours)
//   C: lam  <- lam + grad^T (dbi hbar)   (hbar taken at the -1 neighbours),
//       then lam_z += dL/dtrace(step - 1) at the probes
// g_ca/g_cbi are per-batch-chunk partial planes (gstride floats apart).
struct Adj3DArgs {
  int Mx, My, Mz, pitch, B;
  long long sx, bstride;
  float* lam[3];
  float* hbar[3];
  float* q[3];
  float* ad[3];
  float* al[3];                   // adjoint of J_Lorentz at this reversed step
  float* al_prev[3];              // ... one reversed step back; receives the new al
  const float* ca;
  const float* cb;
  const float* da;                // null -> da = 1, dbi = ch
  const float* dbi;
  const float* d_k1;
  const float* d_k2dt;
  const float* l_k1;
  const float* l_k2dt;
  const float* l_k3;
  float ch, idx;
  const float* e_s;               // E at step s   [3][B][vol]
  const float* e_s1;              // E at step s+1 [3][B][vol]
  long long e_cstride;
  const long long* src_cell;
  const float* wave;
  long long wave_stride_b;
  int step, T, P;
  const float* dtr;               // [B][P][T]
  const short* probe_head;
  const short* probe_next;
  float* g_ca;
  float* g_cbi;
  long long gstride;
};

// A of step a.step (with C of step a.step + 1 folded in when with_c: the
// probe seed it adds is dL/dtrace(a.step)), then B.
hipError_t launch_adj3d(const Adj3DArgs& a, int poles, int bb, bool with_c, hipStream_t s);
// C of step a.step alone (the last reversed step)
hipError_t launch_adj3d_c(const Adj3DArgs& a, int bb, hipStream_t s);
// lam_z[b][probe cell] += dtr[b][q][step] for every TX and probe
hipError_t launch_adj3d_seed(float* lam_z, const long long* probe_cell, const float* dtr,
                             int B, int P, int T, int step, long long bstride, hipStream_t s);

// One fused 3D step (H and E+ADE, inject, probes) from the `*_in` fields to
// the `*_out` fields (ping-pong; ADE currents as in Step3DArgs).  Each E cell
// recomputes the six neighbouring H values it needs from the input fields,
// so H and E are streamed once per step instead of twice.
struct Step3DFArgs {
  Step3DArgs a;                   // geometry, maps, ADE, sources, probes, e_save
  const float* e_in[3];
  const float* h_in[3];
  float* e_out[3];
  float* h_out[3];
  int xcd_order;                  // 1: renumber workgroups so each XCD walks contiguous x-slabs
  int seg_len;                    // marching form: x-slabs per workgroup
};

hipError_t launch_step3d_fused(const Step3DFArgs& a, int poles, int bb, hipStream_t s);
// The same step as a march along x: each thread owns a (y, z) column of
// seg_len slabs and carries E of the current slab and the new hy, hz of the
// previous slab in registers, so no slab is read twice from HBM.
hipError_t launch_step3d_march(const Step3DFArgs& a, int poles, int bb, hipStream_t s);

// Temporally blocked 3D step: K = 1 or 2 full steps per launch (H, E + ADE,
// Ez injection, Ez probes, optional E store per step) on 16 x 64 (y, z)
// windows that march along x.  The state lives in two sets: each field
// component is one allocation of two nbox-float sets, the ADE currents one
// packed array [set][B][cell][nj] (Debye J x, y, z; Lorentz J x, y, z;
// Lorentz J one step older x, y, z; absent poles omitted).  A launch reads
// set in_set and writes the other.
struct TB3DArgs {
  int Mx, My, Mz, pitch, B;
  long long sx, bstride;          // floats per x-slab, per TX volume
  long long nbox;                 // floats per field set (B * bstride)
  int in_set;
  float* e[3]; float* h[3];       // set-0 bases
  float* jp;                      // packed ADE currents (null without poles)
  const float* mp;                // packed maps [cell][map_pack(poles, magnetic).n]
  int magnetic;                   // da / dbi vary per cell (else da = 1, dbi = ch)
  float ch, idx;
  const long long* src_cell;
  const float* wave;
  long long wave_stride_b;
  int step, T, P;                 // step: the launch's first step
  const short* probe_head;
  const short* probe_next;
  float* traces;
  float* e_save;                  // E after the first step ([3][nbox]); the next step's 3 * nbox later
  int tiles_y, tiles_z;           // ceil((My-1)/(16-2K)), ceil((Mz-1)/(64-2K))
  int order;                      // workgroup order: 0 dispatch (y tiles fastest), 1 per XCD, 2 per XCD TX
                                  // fastest, 3 per XCD in groups of tx_group TX, then y tiles
  int tx_group;
  int b0, nb;                     // the launch steps TX [b0, b0 + nb) of the B (arrays span all B)
  // piecewise-constant media without mu_r / sigma_m maps: one byte per cell
  // (box layout, no batch axis) indexing a table of kMat3Floats floats per
  // material, in the order of a packed record (ca, cb, Debye k1, k2/dt,
  // Lorentz k1, k2/dt, k3, absent groups taking no room); null: packed maps
  const unsigned char* mat_id;
  const float* mat_table;
  int n_mat;
};

hipError_t launch_step3d_tb(const TB3DArgs& a, int K, int poles, hipStream_t s);

// Packed 3D maps of the blocked kernels: per cell ca, cb, then Debye k1,
// k2/dt (poles & 1), Lorentz k1, k2/dt, k3 (poles & 2), da, dbi (magnetic);
// absent groups take no room.  Offsets and record length in floats.
struct MapPack {
  int dk, lk, mag, n;
};
// Material-id form: at most kMaxMat3 materials of kMat3Floats table floats
// (two float4 per material; a record of at most 7 maps, the rest padding)
constexpr int kMaxMat3 = 256;
constexpr int kMat3Floats = 8;
constexpr MapPack map_pack(int poles, bool magnetic) {
  return MapPack{2, 2 + ((poles & 1) ? 2 : 0), 2 + ((poles & 1) ? 2 : 0) + ((poles & 2) ? 3 : 0),
                 2 + ((poles & 1) ? 2 : 0) + ((poles & 2) ? 3 : 0) + (magnetic ? 2 : 0)};
}

// Temporally blocked time-reversed 3D step: reversed steps step, step - 1,
// ..., step - K + 1 (K = 1 or 2; each A, B and C of the three-kernel form)
// on 16 x 64 (y, z) windows marching along x.  lam and hbar per component
// are one allocation of two nbox-float sets, the adjoint currents one packed
// array [set][B][cell][nj] like the forward's; a launch reads in_set and
// writes the other.  The gradient sums go to one plane per TX and coefficient.
struct AdjTB3DArgs {
  int Mx, My, Mz, pitch, B;
  long long sx, bstride, nbox;
  int in_set;
  float* lam[3]; float* hbar[3];  // set-0 bases
  float* jp;                      // packed adjoint currents (null without poles)
  const float* mp;                // packed maps as TB3DArgs::mp
  int magnetic;
  float ch, idx;
  const float* e_lo;              // E store slot of step - K + 1 ([3][nbox] per slot), then
                                  // the K slots after it
  const long long* src_cell;
  const float* wave;
  long long wave_stride_b;
  int step, T, P;
  const float* dtr;               // [B][P][T]
  const short* probe_head;
  const short* probe_next;
  float* g;                       // [2][B][gstride]: sums for ca, then for cb (as Adj3DArgs)
  long long gstride;
  int tiles_y, tiles_z;           // ceil((My-1)/(16-2K)), ceil((Mz-1)/(64-2K))
  int order;                      // 0: dispatch order (y tiles fastest), 1: the same with
                                  // contiguous ranges per XCD, 2: per XCD with the TX fastest
  int b0, nb;                     // the launch reverses TX [b0, b0 + nb) of the B (arrays span all B)
  const unsigned char* mat_id;    // material ids and table as TB3DArgs (null: packed maps)
  const float* mat_table;
  int n_mat;
};

hipError_t launch_adj3d_tb(const AdjTB3DArgs& a, int K, int poles, hipStream_t s);

hipError_t launch_h3d(const Step3DArgs& a, int bb, hipStream_t s);
hipError_t launch_e3d(const Step3DArgs& a, int poles, int bb, hipStream_t s);

}  // namespace ceep
