// Launch interface of the 2D step kernels (device pointers, FP32).
#pragma once
#include <hip/hip_runtime.h>

#include "fdtd_common.h"

namespace ceep {

struct Step2DArgs {
  int Mx, My, pitch, B;          // cavity box, row pitch (floats), batch
  long long bstride;             // floats per batch plane (Mx * pitch)
  const float* ez_in;
  const float* hx_in;
  const float* hy_in;
  float* ez_out;
  float* hx_out;
  float* hy_out;
  float* jd;                     // Debye J, updated in place
  float* jl_cur;                 // Lorentz J at the current step
  float* jl_prev;                // Lorentz J one step back; receives J_next
  const float* ca;
  const float* cb;
  const float* d_k1;
  const float* d_k2dt;
  const float* l_k1;
  const float* l_k2dt;
  const float* l_k3;
  float ch;                      // dt / (mu0 dx)
  float idx;                     // 1 / dx
  const int* src_cell;           // [B] a * pitch + c of each TX
  const float* wave;             // [T] or [B][T]
  long long wave_stride_b;       // 0 when one waveform is shared
  int step, T, P;
  const short* probe_head;       // per box cell: first probe id here, or -1
  const short* probe_next;       // [P]: next probe id on the same cell, or -1
  float* traces;                 // [B][P][T]
  float* e_save;                 // if set: E after this step, [B][plane] (adjoint store)
};

// One adjoint (time-reversed) step of the non-dispersive 2D update.  With
// q = cbi * lam_e (cbi = cb / dx, zero off the E interior):
//   mu_x' = mu_x + q(a, c+1) - q(a, c)      mu_y' = mu_y + q(a, c) - q(a+1, c)
//   lam_e' = ca lam_e + ch [(mu_y'(a-1) - mu_y'(a)) - (mu_x'(c-1) - mu_x'(c))]
//            + dL/dtrace(step-1) at the probes
// and, summed over the batch, g_ca += lam_e e^s,
// g_cbi += lam_e (e^{s+1} - ca e^s - source(s)).
struct Adj2DArgs {
  int Mx, My, pitch, B;
  long long bstride;
  const float* le_in;
  const float* mx_in;
  const float* my_in;
  float* le_out;
  float* mx_out;
  float* my_out;
  const float* ca;
  const float* cb;
  float ch, idx;
  const float* e_s;              // forward E at step s   [B][plane]
  const float* e_s1;             // forward E at step s+1 [B][plane]
  const int* src_cell;
  const float* wave;
  long long wave_stride_b;
  int step, T, P;                // step = s (this launch computes lam_e^s)
  const float* dtr;              // [B][P][T] dL/dtrace
  const short* probe_head;
  const short* probe_next;
  float* g_ca;                   // [plane] accumulated over steps and TX
  float* g_cbi;
};

hipError_t launch_adj2d(const Adj2DArgs& a, hipStream_t s);
// lam_e[b][probe q] += dtr[b][q][step] for every TX and probe
hipError_t launch_adj_seed(float* le, const int* probe_cell, const float* dtr, int B, int P,
                           int T, int step, long long bstride, hipStream_t s);

// FP64 fields for the gradient path's forward and recomputation passes
// (non-dispersive; same layout and pitch as the FP32 buffers, in doubles).
struct Step2D64Args {
  int Mx, My, pitch, B;
  long long bstride;
  const double* ez_in;
  const double* hx_in;
  const double* hy_in;
  double* ez_out;
  double* hx_out;
  double* hy_out;
  const double* ca;              // 0 off the E interior
  const double* cb;
  double ch;                     // (dt / mu0) * (1 / dx)
  double idx;                    // 1 / dx
  const int* src_cell;
  const double* wave;            // [T] or [B][T]
  long long wave_stride_b;
  int step, T, P;
  const short* probe_head;
  const short* probe_next;
  double* traces;                // [B][P][T]
  float* traces32;               // the same samples rounded to FP32
  float* e_save;                 // if set: E after this step (FP32) [B][plane]
};

hipError_t launch_step2d_f64(const Step2D64Args& a, hipStream_t s);

// Temporally blocked form of step2d_f64_kernel: K steps per launch on R x 64
// windows split over the four wavefronts of a workgroup (R / 4 rows each,
// halo rows through LDS), one workgroup per (tile, TX chunk).  Bitwise the
// one-step kernel's results.  R in {16, 32}, K in {2, 4}.
struct TB2D64Args {
  int Mx, My, pitch, B;
  long long bstride;
  const double* ez_in; const double* hx_in; const double* hy_in;
  double* ez_out; double* hx_out; double* hy_out;
  const double* ca;              // 0 off the E interior
  const double* cb;
  double ch, idx;
  const int* src_cell;
  const double* wave;
  long long wave_stride_b;
  int step0, ksteps, T, P;
  const int* probe_cell;         // [P] a * pitch + c, or -1 outside the box
  double* traces;                // [B][P][T]
  float* traces32;
  float* e_save;                 // if set: E after local step s (FP32) -> e_save + s * slot
  long long slot;
  int tiles_x, tiles_y, bpw, ntasks, xcd_order;
};

hipError_t launch_step2d_f64_tb(const TB2D64Args& a, int R, int K, hipStream_t s);

// Physical-CPML mode (full grid, the reference single-run solver's step
// order and convolutional PML); profiles per depth index.
struct Cpml2DArgs {
  int Mx, My, pitch, B, t;        // full grid, CPML thickness t
  long long bstride;
  float* ez; float* hx; float* hy;
  float* jd; float* jl_cur; float* jl_prev;
  float* psi_hyx; float* psi_hxy; float* psi_ezx; float* psi_ezy;   // [B][plane]
  const float* ca; const float* cb;
  const float* d_k1; const float* d_k2dt;
  const float* l_k1; const float* l_k2dt; const float* l_k3;
  float dbdy;                     // (dt / mu0) / dx
  float idx;                      // 1 / dx
  float dtmu;                     // dt / mu0
  float dteps;                    // dt / eps0
  float bh[kMaxCpml], chh[kMaxCpml], be[kMaxCpml], ce[kMaxCpml];
  const int* src_cell;
  const float* wave;
  long long wave_stride_b;
  int step, T, P;
  const short* probe_head;
  const short* probe_next;
  float* traces;
};

hipError_t launch_cpml2d(const Cpml2DArgs& a, int poles, hipStream_t s);

// poles: bit 0 = Debye pole present, bit 1 = Lorentz pole present.
// bb: batch planes per thread (1, 2, 4, 8).
hipError_t launch_step2d(const Step2DArgs& a, int poles, int bb, hipStream_t s);

}  // namespace ceep

namespace ceep {

// Temporal blocking: one wavefront holds an R x 64 window (rows = x, lanes =
// y) of the state in registers and advances it ksteps <= K steps; the output
// window (rows [K, R-K), lanes [K, 64-K)) is written back.  All state,
// including the Debye current, is ping-ponged.
struct TB2DArgs {
  int Mx, My, pitch, B;
  long long bstride;
  const float* ez_in;
  const float* hx_in;
  const float* hy_in;
  const float* jd_in;
  float* ez_out;
  float* hx_out;
  float* hy_out;
  float* jd_out;
  const float* ca;                // zero outside the E interior
  const float* cb;
  const float* d_k1;
  const float* d_k2dt;
  float ch, idx;
  const int* src_cell;            // [B] a * pitch + c
  const float* wave;
  long long wave_stride_b;
  int step0, ksteps, T, P;
  const int* probe_cell;          // [P] a * pitch + c, or -1 outside the box
  float* traces;
  int tiles_x, tiles_y;           // output tiles over rows / columns
  int bpw;                        // batch planes per wavefront
  int ntasks;                     // tiles_x * tiles_y * ceil(B / bpw)
  float* e_save;                  // if set: E after local step s -> e_save + s * slot ([B][plane])
  long long slot;                 // floats per e_save slot
  int xcd_order;                  // contiguous tile ranges per XCD (xcd_block)
  // difference form (poles == kPolesDiff): the state is E - E_ref of a
  // reference run with coefficients ca_r, cb_r; its E^s sits at
  // e_ref + s * ref_slot ([B][plane]).  The E step adds
  // d_a E_ref^s + d_b (E_ref^{s+1} - source^s) with d_a = ca - cb ca_r / cb_r,
  // d_b = cb / cb_r - 1 (zero where the two media agree), only in the tiles
  // diff_tile flags; no source is injected (it cancels).
  const float* d_a;
  const float* d_b;
  const float* e_ref;
  long long ref_slot;
  const unsigned char* diff_tile; // [tiles_x * tiles_y]
  // piecewise-constant media: per-cell material id ([Mx][pitch] bytes; id 0 =
  // all coefficients 0, off the E interior) and the table of (ca, cb, k1,
  // k2dt) per id; null: the full maps above are read
  const unsigned char* mat_id;
  const float4* mat_table;
  int n_mat;                      // <= kMaxMaterials
  // optional subset of the output tiles (four- and eight-wavefront forward):
  // task = chunk * n_list + i steps tile tile_list[i]; null: every tile
  const int* tile_list;
  int n_list;
  int no_inject;                  // 1: no source term (the difference form's
                                  // tiles where the media agree)
};

constexpr int kMaxMaterials = 256;

// The difference form's `poles` value (non-dispersive media only).
constexpr int kPolesDiff = 4;

// poles: 0 or 1 (Debye); R in {16, 24, 32}; K in {1, 2, 3, 4}.
hipError_t launch_step2d_tb(const TB2DArgs& a, int poles, int R, int K, hipStream_t s);
// The same update with an R x 64 window split over the four wavefronts of a
// workgroup (R / 4 rows each, halo rows through LDS); one workgroup per task.
// R in {32, 48, 64}, K in {4, 6, 8}, poles 0, 1 or kPolesDiff.
// NW = 4 or 8 wavefronts per window; step2d_mw_shape_ok lists the (R, K)
// instantiated for each
bool step2d_mw_shape_ok(int R, int K, int NW);
hipError_t launch_step2d_mw(const TB2DArgs& a, int poles, int R, int K, int NW, hipStream_t s);

// Temporally blocked adjoint: K adjoint steps (step_hi-1 down to step_hi-ksteps)
// per launch on R x 64 register windows, the transpose of step2d_tb_kernel's
// update; gradient sums go to one partial plane per TX chunk (no atomics).
struct AdjTB2DArgs {
  int Mx, My, pitch, B;
  long long bstride;
  const float* le_in;
  const float* mx_in;
  const float* my_in;
  float* le_out;
  float* mx_out;
  float* my_out;
  const float* ca;
  const float* cb;
  float ch, idx;
  const float* e_store;           // forward E by slot: E^s at e_store + (s - seg0) * slot
  long long slot;
  int seg0;
  const int* src_cell;
  const float* wave;
  long long wave_stride_b;
  int step_hi, ksteps, T, P;
  const float* dtr;               // [B][P][T] dL/dtrace
  const int* probe_cell;
  float* g_ca;                    // [nchunks][gstride] partial sums, one plane per TX chunk
  float* g_cbi;
  long long gstride;
  int tiles_x, tiles_y, bpw, ntasks;
  // reversible form: the forward state at step_hi (ez, hx, hy) in, and at
  // step_hi - ksteps out; e_store is unused
  const float* fe_in;
  const float* fhx_in;
  const float* fhy_in;
  float* fe_out;
  float* fhx_out;
  float* fhy_out;
  int xcd_order;                  // contiguous tile ranges per XCD (xcd_block)
};

hipError_t launch_adj2d_tb(const AdjTB2DArgs& a, int R, int K, hipStream_t s);
// The reversible blocked adjoint: the forward state is stepped backward
// (E^s = (E^{s+1} - source - cb/dx curl H^{s+1/2}) / ca, then
// H^{s-1/2} = H^{s+1/2} -+ ch dE^s) inside the same windows as the adjoint,
// so no segment is recomputed and no E store is read.
hipError_t launch_adj2d_rev_tb(const AdjTB2DArgs& a, int R, int K, hipStream_t s);
// The same reversed steps with an R x 64 window split over the NW
// wavefronts of a workgroup (R / NW rows each; halo rows through LDS); one
// workgroup per task.  NW = 4: R in {16, 24, 32}; NW = 8: R in {32, 48}.
// prefetch: the next TX's planes are loaded while the current one is stepped
// (NW = 4: R = 32, K = 6; NW = 8: R = 32, K = 4..6)
bool adj2d_rev_mw_shape_ok(int R, int K, int NW, bool prefetch);
hipError_t launch_adj2d_rev_mw(const AdjTB2DArgs& a, int R, int K, int NW, bool prefetch,
                               hipStream_t s);

// One problem of an uneven batch: its blocked-kernel arguments (buffers set
// per round from the ping-pong pairs) and the number of K-step rounds.
struct MultiDesc {
  TB2DArgs a;
  float* f[2][3];
  float* jd[2];
  int nrounds;
};

// Uneven batches: items are (round, problem, task), enumerated as segments
// of one problem's tasks in one round; one launch per round.
struct MultiArgs {
  const MultiDesc* probs;
  int nprob;
  const long long* seg_start;     // [nseg] first item of each segment (ascending)
  const int* seg_prob;
  const int* seg_round;
  int nseg;
  unsigned long long total;       // number of items
  int xcd_order;                  // contiguous item ranges per XCD
};

// One persistent launch of an uneven batch: items are whole (problem, TX
// chunk) runs, taken by workgroups of four-wavefront groups from a counter
// over `items` (largest first); see persist_mw_kernel.
struct PersistItem {
  int prob;
  int chunk;                      // TX planes [chunk * bpw, min((chunk + 1) * bpw, B))
};
struct PersistArgs {
  const MultiDesc* probs;
  const PersistItem* items;
  int nitems;
  unsigned* queue;                // next item (zeroed before the launch)
};
// Four-wavefront tasks (R = 32 or 48, K = 4 or 6), `blocks` workgroups of
// persist_threads(poles, R) threads (groups of four wavefronts).
hipError_t launch_persist_mw(const PersistArgs& m, int poles, int R, int K, bool mat_ids,
                             int blocks, hipStream_t s);
int persist_threads(int poles, int R);
int persist_mw_blocks_per_cu(int poles, int R, int K, bool mat_ids);

// One round's items [first, first + count) (segments [seg0, seg1)) as one launch.
hipError_t launch_round_tb(const MultiArgs& m, int poles, int R, int K, int seg0, int seg1,
                           long long first, long long count, hipStream_t s);
// The same rounds with NW wavefronts per window (NW = 4: R = 32 or 48 window
// rows, K = 4 or 6; NW = 8: R = 64, K = 6): one workgroup per work item.
hipError_t launch_round_mw(const MultiArgs& m, int poles, int R, int K, int NW, int seg0, int seg1,
                           long long first, long long count, bool mat_ids, hipStream_t s);
// dst[i] = sum over k < nparts of src[k * stride + i], i < n
hipError_t launch_sum_parts(const float* src, int nparts, long long stride, long long n,
                            float* dst, hipStream_t s);

}  // namespace ceep
