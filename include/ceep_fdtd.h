/*
 * ceep_fdtd.h -- C ABI of the MI355X batched FDTD engine (libceep_fdtd.so).
 *
 * Plain pointers and sizes only.  Host arrays are C-order float64 unless
 * stated; positions are int64 grid indices.  Every call returns a status code
 * (CEEP_OK = 0); on failure ceep_fdtd_last_error() returns a message that
 * stays valid until the calling thread's next failing call.
 *
 * The reference has no FFI: its boundary is the Python class API of
 * BatchedFDTD2D / BatchedFDTD3D (src/ceep/solvers/fdtd_2d_batched.py,
 * src/ceep/solvers/fdtd_3d_batched.py).  Each entry point names the reference
 * method whose work it takes over.  A handle owns one HIP stream and its
 * device buffers; calls on one handle must not overlap, distinct handles are
 * independent.
 */
#ifndef CEEP_FDTD_H
#define CEEP_FDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ceep_fdtd ceep_fdtd;

enum {
  CEEP_OK = 0,
  CEEP_EINVAL = 1,       /* bad argument (ValueError / IndexError in Python) */
  CEEP_EHIP = 2,         /* HIP runtime error */
  CEEP_ENOMEM = 3,       /* device allocation failed */
  CEEP_ESTATE = 4,       /* call out of order (e.g. DFT before run) */
  CEEP_EUNSUPPORTED = 5  /* valid input this engine does not handle */
};

/* Box modes for ceep_fdtd_create. */
enum {
  CEEP_BOX_REFERENCE = 0, /* BatchedFDTD.run(): E interior [n+1, N-n-1], box [n, N-n] */
  CEEP_BOX_FULL = 1,      /* BatchedFDTD2D.run_cpu(): PEC box over the full grid */
  CEEP_BOX_CPML = 2       /* opt-in physical absorbing boundary: the full grid stepped in
                             the order of FDTDBase.step (fdtd_base.py:161-185) with the CPML
                             class's psi recursions on slabs of cpml_n cells (2D:
                             absorbing.py:266-561; 3D: :567-943 around the update of
                             fdtd_3d.py:97-171); sources and probes anywhere */
};

/* Message of the last failing call on this thread ("" if none). */
const char* ceep_fdtd_last_error(void);

/* ABI revision of this header (bumped on incompatible changes). */
int ceep_fdtd_abi_version(void);

/* Replaces BatchedFDTD2D.__init__ + _build (fdtd_2d_batched.py:66-98, 138-193)
 * and BatchedFDTD3D.__init__ + _build (fdtd_3d_batched.py:104-141, 266-326):
 * allocates device state for a batch of `batch` TX and `n_probes` RX over
 * `total_steps` steps.  ndim = 2 (shape = {nx, ny}) or 3 (shape = {nx, ny, nz});
 * dt is supplied by the caller (reference: dx/(c sqrt 2)*0.99 in 2D,
 * cfl_timestep_3d(dx,dx,dx,0.5)*0.99 in 3D). */
int ceep_fdtd_create(int ndim, const int64_t* shape, int cpml_n, int box_mode, int batch,
                     int n_probes, int total_steps, double dx, double dt, ceep_fdtd** out);

int ceep_fdtd_destroy(ceep_fdtd* h);

/* Replaces the coefficient part of _build (fdtd_2d_batched.py:152-165;
 * fdtd_3d_batched.py:282-296): full-grid maps eps_r, sigma_e (and in 3D
 * optional mu_r, sigma_m; NULL = 1 and 0). ca/cb/da/db are formed in FP64
 * on the host, including eps_eff = eps_r*eps0 + sum(k2/2) of any ADE poles
 * (core/grid.py:178-184), and stored as FP32 on the device. */
int ceep_fdtd_set_materials(ceep_fdtd* h, const double* eps_r, const double* sigma_e,
                            const double* mu_r, const double* sigma_m);

/* One Debye pole per cell (DispersiveManager Debye branch,
 * materials/dispersive.py:253-261): full-grid delta_eps and tau maps, tau <= 0
 * meaning no pole.  With poles set, eps_r of set_materials is eps_inf. */
int ceep_fdtd_set_debye(ceep_fdtd* h, const double* delta_eps, const double* tau);

/* One Lorentz pole per cell (dispersive.py:278-299): full-grid delta_eps,
 * omega_0 (rad/s; <= 0 meaning no pole) and delta (rad/s) maps. */
int ceep_fdtd_set_lorentz(ceep_fdtd* h, const double* delta_eps, const double* omega_0,
                          const double* delta);

/* Remove all ADE poles. */
int ceep_fdtd_clear_poles(ceep_fdtd* h);

/* Source positions (batch x ndim) and waveform: per_batch = 0 -> one (T,)
 * waveform for every TX (fdtd_2d_batched.py:509-515), 1 -> (batch, T). */
int ceep_fdtd_set_sources(ceep_fdtd* h, const int64_t* pos, const double* waveform,
                          int per_batch);

/* Probe positions (n_probes x ndim) recorded for every TX (:517-526). */
int ceep_fdtd_set_probes(ceep_fdtd* h, const int64_t* pos);

/* Zero every field and ADE current (the state _build allocates,
 * fdtd_2d_batched.py:147-149, fdtd_3d_batched.py:274-279).  A new handle
 * starts zeroed. */
int ceep_fdtd_reset(ceep_fdtd* h);

/* Replaces the time loop of run() (fdtd_2d_batched.py:454-526,
 * fdtd_3d_batched.py:918-958): advances total_steps steps on the device from
 * the current state (like the reference, a second run continues from the
 * fields the first one left) and, if traces_out != NULL, copies the (batch,
 * n_probes, total_steps) traces to it as float64 (the probe_data layout of
 * fdtd_2d_batched.py:188-191). */
int ceep_fdtd_run(ceep_fdtd* h, double* traces_out);

/* Copy the traces of the last run (batch, n_probes, total_steps) float64. */
int ceep_fdtd_get_traces(ceep_fdtd* h, double* traces_out);

/* Running DFT of the last run's traces (DFTMonitor.update, solvers/dft.py:55-90,
 * summed over every step) into the MultistaticSParameters layout
 * (s_params.py:99): s_out is (n_freq, n_probes, batch) complex as interleaved
 * float64 (re, im). */
int ceep_fdtd_dft(ceep_fdtd* h, const double* freqs, int n_freq, double* s_out);

/* ceep_fdtd_dft without any device or host allocation: the frequency list, the
 * twiddle table (formed on the device from the same FP64 expressions) and the
 * S staging live in one static device array of 16 MiB per GPU, which every
 * GPU the library runs on keeps until the process ends.  Rejects (CEEP_EINVAL)
 * a call whose n_freq x total_steps twiddles and S staging exceed it.  All
 * calls on one GPU share that array: calls that overlap on one GPU (other
 * threads or streams) race.  No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
int ceep_fdtd_dft_noalloc(ceep_fdtd* h, const double* freqs, int n_freq, double* s_out);

/* Copy the current field component comp (2D: 0 = ez, 1 = hx, 2 = hy; 3D:
 * 0..5 = ex, ey, ez, hx, hy, hz) to a full-grid (batch, nx, ny[, nz])
 * float64 array (cells outside the stepped box are exactly 0 in the
 * reference and are written as 0). */
int ceep_fdtd_get_field(ceep_fdtd* h, int comp, double* out);

/* BatchedFDTD3D.compute_energy (fdtd_3d_batched.py:987-1011): sum over
 * the batch and grid of (eps0/2 |E|^2 + mu0/2 |H|^2) dV; 2D: dV = dx*dy. */
int ceep_fdtd_energy(ceep_fdtd* h, double* out);

/* Timing: `reps` device-resident runs, each from a zeroed state (no host
 * copies), bracketed by HIP events on the handle's stream; *ms_per_run =
 * mean elapsed ms per run.  Leaves the state of the last run. */
int ceep_fdtd_time_runs(ceep_fdtd* h, int reps, double* ms_per_run);

/* Forward run for the adjoint (2D, non-dispersive): zeroes the state, runs
 * total_steps steps like ceep_fdtd_run and keeps the full state every `every`
 * steps as a checkpoint.  traces_out as in ceep_fdtd_run.  No counterpart in
 * the reference (it has no adjoint; src/ceep/ai/__init__.py is a docstring). */
int ceep_fdtd_run_checkpointed(ceep_fdtd* h, int every, double* traces_out);

/* Discrete adjoint of the last checkpointed run.  dtraces = dL/dtrace,
 * (batch, n_probes, total_steps) float64, for any loss L of the traces; on
 * return grad_eps_r and grad_sigma (full grid, float64) hold dL/d eps_r and
 * dL/d sigma per cell (0 outside the stepped E interior, where the reference's
 * result does not depend on them).  Each checkpoint segment is recomputed
 * forward (storing E) and then stepped backward; the last segment's E was
 * stored by the checkpointed run itself, so with every >= total_steps nothing
 * is recomputed.  Leaves the fields zeroed; the checkpoints stay valid for
 * further adjoint calls. */
int ceep_fdtd_adjoint(ceep_fdtd* h, const double* dtraces, double* grad_eps_r,
                      double* grad_sigma);

/* Reference medium of the difference form (2D, cavity modes): eps_r and
 * sigma_e (full grid, float64) of the medium whose traces the misfit is
 * measured against -- e.g. the target phantom of L = sum |S - S_target|^2.
 * Both NULL clears it.  The next ceep_fdtd_run_residual steps that medium once
 * with this handle's sources, probes and launch shape, keeping its E after
 * every step ((total_steps + 1) x batch x box plane floats of device memory)
 * and its state at every checkpoint.  No counterpart in the reference (it has
 * no gradient path); the target is what its S-matrix workflow compares
 * against (src/ceep/solvers/s_params.py:186-231). */
int ceep_fdtd_set_reference(ceep_fdtd* h, const double* eps_r, const double* sigma_e);

/* Checkpointed forward of the difference E - E_ref (FP32, 2D non-dispersive,
 * four-wavefront blocked forward, reversible adjoint): the state is the
 * difference itself, stepped with this medium's coefficients plus the source
 * term d_a E_ref^s + d_b (E_ref^{s+1} - source^s) where the media differ, so
 * FP32 rounds the residual rather than two nearly equal totals.  residual_out
 * (batch, n_probes, total_steps) gets trace(eps) - trace(eps_ref) (or NULL);
 * reference_out the reference run's traces (or NULL).  The checkpoints it
 * leaves hold the medium's total state, so ceep_fdtd_adjoint runs after it as
 * after ceep_fdtd_run_checkpointed.  No counterpart in the reference. */
int ceep_fdtd_run_residual(ceep_fdtd* h, int every, double* residual_out, double* reference_out);

/* The whole gradient of L = sum over (f, rx, tx) of |S - S_target|^2 at
 * n_freq frequencies (2D, non-dispersive, FP32 forward) in one call, with
 * no trace or dL/dtrace round trip through the host: the checkpointed forward
 * (s_target = (n_freq, n_probes, batch) complex as interleaved float64 pairs:
 * ceep_fdtd_run_checkpointed; s_target NULL with a reference medium set:
 * ceep_fdtd_run_residual, whose residual is S - S_target itself), the
 * S-matrix of the traces, the loss and
 * dL/dtrace on the device in FP64 sums, then the adjoint as ceep_fdtd_adjoint.
 * loss_out gets L; grad_eps_r and grad_sigma (full grid, float64) dL/d eps_r
 * and dL/d sigma.  No counterpart in the reference (no gradient path; the
 * S-matrix is that of src/ceep/solvers/dft.py:75-90, s_params.py:70-99). */
int ceep_fdtd_gradient_misfit(ceep_fdtd* h, int every, const double* freqs, int n_freq,
                              const double* s_target, double* loss_out, double* grad_eps_r,
                              double* grad_sigma);

/* The reference medium's traces (batch, n_probes, total_steps) float64 from
 * its last run (ceep_fdtd_run_residual's reference_out). */
int ceep_fdtd_get_reference_traces(ceep_fdtd* h, double* out);

/* Gradient-path storage: out[0] = E-store slots (0 when there is none),
 * out[1] = of them beyond one segment (3D, tuning key 25), out[2] = checkpoints. */
int ceep_fdtd_store_info(ceep_fdtd* h, int64_t* out);

/* Tuning knobs (keys 0-4 give the same results to FP32 rounding for every setting):
 *   key 0: batch planes per thread of the one-step kernels (1, 2, 4, 8; the
 *          3D one-step kernels use at most 4, the 3D adjoint kernels all four)
 *   key 1: replay the time loop as a captured hipGraph (0/1)
 *   key 2: 2D temporal depth K = steps per launch (0 = one-step kernel; 1..4
 *          with one wavefront per window, 4, 6 or 8 with four, key 16)
 *   key 3: 2D window rows of the blocked kernel (16, 24, 32 with one
 *          wavefront per window; 32, 48, 64 with four)
 *   key 4: 2D batch planes per window of the blocked kernel (>= 1)
 *   key 5: precision of the gradient path's forward pass (0 = FP32, 1 =
 *          FP64; 2D non-dispersive): with 1, ceep_fdtd_run_checkpointed steps
 *          FP64 fields and returns FP64 traces, for misfits whose residual
 *          lies below FP32 trace accuracy.  The adjoint stays FP32: the
 *          reversible form (key 7) starts from FP32 roundings of the FP64
 *          segment-end states, the recompute form re-steps each segment in
 *          FP64.  Not a launch-shape knob: results differ.
 *   key 6: 3D step form: 4 (default) = temporally blocked, key-17 steps
 *          per launch on 16 x 64 (y, z) windows marching along x, the whole
 *          state in two ping-pong sets; 0 = H kernel then E kernel in place;
 *          1 = one fused H+E launch per step on ping-pong fields, workgroups
 *          ordered so each XCD walks contiguous x-slabs; 2 = fused, dispatch
 *          order; 3 = fused, each thread marching along x over key-8 slabs.
 *          All forms give bit-identical results.  Switching between 0, 1-3
 *          and 4 (different state layouts) clears the fields.  Form 4 steps
 *          with form 0 where a TX's two state sets exceed 32-bit byte
 *          offsets ((B + 1) x box volume x 4 B >= 2^31, or box volume x 36 B
 *          >= 2^31).
 *   key 7: 2D blocked adjoint form: 1 = reversible (the forward state is
 *          rebuilt backward inside the adjoint launches from the checkpoint at
 *          each segment's end; no recomputation, no E store; default),
 *          0 = recompute each segment into an E store.
 *   key 8: x-slabs per workgroup of the marching 3D form (default 24).
 *   key 9: 2D blocked kernels: 1 = renumber workgroups so each XCD runs a
 *          contiguous range of tiles (neighbouring tiles share an L2; default),
 *          0 = dispatch order.  Results are identical.
 *   key 10, 11: window rows (16, 24, 32, 48; default 32 for the multi-
 *          wavefront reversible kernel, 16 otherwise) and batch planes per window
 *          (default 4) of the 2D blocked adjoint, which holds twice the
 *          forward's state per cell and prefers other shapes than the forward
 *          (keys 3, 4).
 *   key 14: 2D reversible adjoint: wavefronts per window, 1, 4 or 8
 *          (default 4: each wavefront holds a quarter of the window's rows,
 *          halo rows pass through LDS; 8 takes 32- or 48-row windows).
 *   key 15: 2D blocked adjoint: reversed steps per launch, 1..6 (5 and 6
 *          only with the multi-wavefront reversible kernel, 6 only on 32- or
 *          48-row windows; default 6 there, 5 on smaller windows of that
 *          kernel, 4 otherwise; eight wavefronts take 4..6).
 *   key 16: 2D blocked forward: wavefronts per window, 1 (default), 4 or 8
 *          (eight: 32 rows at depth 4, 48 at 4 or 6, 64 at 4, 6 or 8).
 *          Results are bit-identical between the forms.
 *   key 12: 3D workgroup order: 0 = dispatch order (default; blocked form:
 *          y tiles fastest, then z tiles, then TX), 1 = the same renumbered so
 *          each XCD runs a contiguous range, 2 = per XCD with the TX (batch
 *          chunk) fastest, 3 = blocked form only: per XCD in groups of key-18
 *          TX, then y tiles.  Results are identical; 0 measured fastest.
 *   key 17: 3D blocked form: steps per launch, 1 or 2 (default 2).
 *   key 18: 3D blocked form, order 3: TX per group (default 4).
 *   key 19: 3D gradient with the blocked forward (key 6 = 4): reversed steps
 *          per adjoint launch, 1 or 2 (default 2; one launch runs A, B and C
 *          of each of its steps on 16 x 64 windows marching along x), or 0 =
 *          three kernels per reversed step.  The adjoint fields are
 *          bit-identical between the forms; the blocked launches sum the
 *          gradient per TX and add the TX at the end, the three-kernel order
 *          when key 0 is 1.  The three kernels run instead where
 *          (2 x B + 1) x box volume x 4 B reaches 2^31.
 *   key 20: 3D blocked adjoint: workgroup order, 0 = dispatch order (y tiles
 *          fastest), 1 = the same renumbered so each XCD runs a contiguous
 *          range (default; y-neighbour tiles share halo rows in an L2),
 *          2 = per XCD with the TX fastest.  Results are identical.
 *   key 22: material ids for piecewise-constant media (at most 256 distinct
 *          coefficient records in the box): one byte per cell plus a table
 *          instead of the per-cell coefficient maps.  0 = off, 1 (default) =
 *          the 2D blocked forward, residual and uneven-round kernels and the
 *          3D blocked adjoint, 2 = also the 3D blocked forward.  Media with
 *          more records, Lorentz poles in 2D, mu_r / sigma_m maps in 3D and
 *          the physical-CPML mode use the maps.  Results are bit-identical.
 *   key 23: difference form (ceep_fdtd_run_residual), tile order of each
 *          blocked step: 0 = the difference kernel over every tile in
 *          order; 1 (default) = the same with the tiles whose window meets
 *          cells where the media differ dispatched first; 2 = those tiles
 *          with the difference kernel on a second stream and the others
 *          with the plain kernel (no source term), joined by events each
 *          step; 3 = the same two launches one after the other.  Results
 *          are identical up to the plain kernel's rounding of the update.
 *   key 24: 2D reversible adjoint: 1 = load the next TX's six planes while
 *          the current one is stepped (four wavefronts: 32 rows, depth 6;
 *          eight: 32 rows, depth 4-6); 0 (default) = load each TX when it
 *          starts.  Results are identical.
 *   key 27: TX groups on streams: 1, 2 (default) or 4.  The 2D blocked
 *          forward (also the difference form and the reference run) and the
 *          reversible 2D adjoint split the batch into that many groups of
 *          whole TX chunks, each stepped on its own stream; the streams fork
 *          once before a run (or a checkpoint segment) and join after it, so
 *          one group's launches fill the partly idle last rounds of the
 *          others'.  Results are bit-identical.
 *   key 28: the same for the 3D blocked forward, groups of whole TX: 1
 *          (default), 2 or 4.  Bit-identical; measured slower (200^3 x 16
 *          TX: 2130 against 2055 ms per run).
 *   key 29: the same for the 3D blocked adjoint: 1 (default), 2 or 4.
 *          Bit-identical; measured slower (the 200^3 x 16 TX gradient:
 *          6.61 and 6.49 s on 2 and 4 streams against 6.24 s on one).
 *   key 25: 3D gradient (blocked adjoint): extra E-store slots beyond one
 *          segment.  The checkpointed forward also keeps E of the last
 *          `extra` steps of the segment before the last, so that segment's
 *          adjoint recomputes only its first every - extra steps.  -1
 *          (default) = as many as the free device memory holds after what the
 *          adjoint still allocates and a 4 GiB margin (at most every - 1),
 *          >= 0 = that many.  Results are identical. */
int ceep_fdtd_set_tuning(ceep_fdtd* h, int key, int value);

/* CPML profile parameters of a CEEP_BOX_CPML handle (defaults: order 4,
 * sigma_factor 2.0, alpha_max 0.05, kappa_max 1.0, as CPML's fields,
 * absorbing.py:226-230).  Replaces: the CPML(...) boundary object handed to
 * FDTD2D / FDTD3D (solvers/fdtd_2d.py, fdtd_3d.py, boundaries/absorbing.py:
 * 174-311, 567-626). */
int ceep_fdtd_set_cpml(ceep_fdtd* h, int order, double sigma_factor, double alpha_max,
                       double kappa_max);

/* Delay-and-sum / delay-multiply-and-sum image from a multistatic S-matrix.
 * Replaces: DelayAndSumBeamformer(s_parameters, antenna_positions,
 * ImagingRegion(x_range, y_range, dx, background_permittivity),
 * frequency_range).reconstruct(method, normalize)
 * (src/ceep/imaging/beamforming.py:86-313).
 * s_matrix: [n_freq][n_ant (RX)][n_ant (TX)] complex128 as interleaved
 * (re, im) doubles (the MultistaticSParameters layout); antennas: [n_ant][2]
 * grid indices; x_range, y_range: [2] pixel index ranges (end excluded);
 * frequency_range: [2] inclusive (f_min, f_max) or NULL for all; method:
 * 0 = DAS, 1 = DMAS; image_out: [nx][ny] magnitude, normalised to [0, 1]
 * (minus the minimum, over the maximum) when normalize != 0.  Synchronous;
 * allocates and frees its own device buffers. */
int ceep_beamform(const double* s_matrix, int n_freq, int n_ant, const double* freqs,
                  const int64_t* antennas, const int64_t* x_range, const int64_t* y_range,
                  double dx, double background_permittivity, const double* frequency_range,
                  int method, int normalize, double* image_out);

/* Geometry of the stepped box: out[0..2] = box size per axis, out[3] =
 * row pitch in floats, out[4] = box origin (grid index of box cell 0). */
int ceep_fdtd_box_info(ceep_fdtd* h, int64_t* out);

/* ---- uneven batches (no counterpart in the reference) ----------------------
 * Many 2D problems of different grid sizes and step counts, each a fully set
 * up handle (materials, sources, probes; no Lorentz poles; one pole set for
 * all), advanced together in one persistent launch or round by round, the
 * clears and the launches captured as one hipGraph on the first run.
 * Traces land in each handle (ceep_fdtd_get_traces, ceep_fdtd_dft). */
typedef struct ceep_fdtd_multi ceep_fdtd_multi;

/* waves = 1: one wavefront per window, K = steps per round (1..4), R =
 * window rows (16, 24, 32); waves = 4: four wavefronts per window (halo rows
 * through LDS), K = 4 or 6, R = 32 or 48; waves = 8: eight, K = 6 or 8, R = 64
 * (per-round mode only).  bpw = TX planes per window.  The
 * handles must outlive the group. */
int ceep_fdtd_multi_create(int n, ceep_fdtd* const* handles, int K, int R, int bpw, int waves,
                           ceep_fdtd_multi** out);

/* Zero every problem and run all of them to their own total_steps, as a
 * captured graph (captured on the first call).  mode 0 (four-wavefront
 * shapes; CEEP_EUNSUPPORTED otherwise): one persistent launch of 1024-thread
 * workgroups, as many as fit at once, each taking whole (problem, TX chunk)
 * runs from a counter over a largest-first list and stepping every round of
 * one alone (its four 4-wavefront groups split each round's windows; a
 * workgroup barrier orders rounds; no workgroup waits for another).  mode 1:
 * every round (K steps of every problem still running) is one launch over
 * all problems' windows, or, with tuning key 27 of the first handle = G > 1,
 * G launches: group g takes TX chunks [c * g / G, c * (g + 1) / G) of every
 * problem of c chunks of bpw planes, and each group's rounds run on a stream
 * of their own (parallel branches of the graph, rejoined at its end).
 * Either mode gives every problem the bits of its own blocked run with the
 * same window shape.  mode 2 (every problem's own run on concurrent streams)
 * is switched off and returns CEEP_EUNSUPPORTED. */
int ceep_fdtd_multi_run(ceep_fdtd_multi* m, int mode);

/* `reps` graph replays of `mode` bracketed by HIP events; *ms = mean ms per replay. */
int ceep_fdtd_multi_time(ceep_fdtd_multi* m, int mode, int reps, double* ms);

/* out[0] = per-round work items, out[1] = (round, problem) segments (both
 * summed over the TX groups of mode 1),
 * out[2] = workgroups of the persistent launch (0 for one-wavefront shapes),
 * out[3] = its (problem, TX chunk) items. */
int ceep_fdtd_multi_info(ceep_fdtd_multi* m, long long* out);

int ceep_fdtd_multi_destroy(ceep_fdtd_multi* m);

#ifdef __cplusplus
}
#endif

#endif /* CEEP_FDTD_H */
