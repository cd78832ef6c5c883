// Host side of libceep_fdtd.so: the C ABI declared in include/ceep_fdtd.h.
//
// A handle holds the batched solver's device state for one geometry: field
// buffers over the stepped box, FP32 coefficient maps, ADE currents, source
// and probe tables and the (batch, probes, steps) trace buffer.  The time loop
// is a sequence of kernel launches on the handle's stream, captured once as a
// hipGraph and replayed for every run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <array>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <string>
#include <vector>

#include "ceep_fdtd.h"
#include "fdtd_common.h"
#include "fdtd2d_kernels.h"
#include "fdtd3d_kernels.h"
#include "reduce_kernels.h"
#include "misfit.h"
#include "imaging.h"
#include "coef2d.h"
#include "limits2d.h"

namespace {

constexpr double kEps0 = 8.854187817e-12;            // core/constants.py:24
constexpr double kMu0 = 4.0 * 3.14159265358979323846 * 1e-7;  // core/constants.py:27
constexpr double kC0 = 299792458.0;                    // core/constants.py:21

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Pinned host staging buffer of at least n floats (grown, never shrunk):
// host<->device copies of it run at full PCIe rate and truly asynchronously.
int pinned(float** p, size_t* cap, size_t n) {
  if (*cap >= n) return CEEP_OK;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipHostMalloc(reinterpret_cast<void**>(p), n * sizeof(float), hipHostMallocDefault) != hipSuccess)
    return fail(CEEP_ENOMEM, "pinned host staging allocation failed");
  *cap = n;
  return CEEP_OK;
}

#define CEEP_HIP(call)                                                               \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(CEEP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

template <class T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return CEEP_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(CEEP_ENOMEM, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                 " B): " + hipGetErrorString(e));
  }
  return CEEP_OK;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

// Scratch of the no-allocation DFT (ceep_fdtd_dft_noalloc): the frequency
// list, the twiddle table and the S staging of one call, in one file-scope
// device array.  Every GPU the library runs on holds its own copy for the
// life of the process; host code reaches it only through the per-device
// address hipGetSymbolAddress returns (scratch_address), never through the
// array's name.  All calls on one GPU share it, and it keeps the last call's
// contents.
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};

// The current device's copy of `scratch`, resolved once per device.
int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(CEEP_EHIP, "no usable current HIP device for the DFT scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CEEP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return CEEP_OK;
}
}  // namespace

struct ceep_fdtd {
  int ndim = 2, box_mode = 0, n = 0, B = 0, P = 0, T = 0;
  unsigned char* scratch = nullptr;   // this handle's GPU's copy of the DFT scratch
  int64_t N[3] = {1, 1, 1};
  double dx = 0, dt = 0;
  int org = 0;                 // grid index of box cell 0 (same on every axis)
  int M[3] = {1, 1, 1};        // box size per axis
  int pitch = 0;               // floats per innermost row
  long long row_block = 0;     // 3D: floats per x-slab (M[1] * pitch)
  long long plane = 0;         // floats per batch plane/volume
  long long full_cells = 0;    // cells of the full grid

  // host-side inputs
  std::vector<double> eps_r, sigma_e, mu_r, sigma_m;
  bool magnetic = false;
  std::vector<double> deb_de, deb_tau, lor_de, lor_w0, lor_dl;
  bool has_debye = false, has_lorentz = false;
  std::vector<int64_t> src, prb;
  bool have_sources = false, have_probes = false;

  // device state
  float* f[2][6] = {{nullptr}};
  float* jd[3] = {nullptr, nullptr, nullptr};
  float* jl[2][3] = {{nullptr}};
  float *ca = nullptr, *cb = nullptr, *da = nullptr, *dbi = nullptr;
  float *dk1 = nullptr, *dk2 = nullptr, *lk1 = nullptr, *lk2 = nullptr, *lk3 = nullptr;
  // 3D blocked form: per-cell packed maps [cell][9] (ca, cb, dk1, dk2dt, lk1,
  // lk2dt, lk3, da, dbi) and ADE currents [set][B][cell][nj] (Debye J x, y, z;
  // Lorentz J x, y, z; Lorentz J one step older x, y, z; absent poles omitted)
  float* mpack = nullptr;
  float* jpack = nullptr;
  size_t mpack_n = 0, jpack_n = 0;   // their float counts (kept across material changes)
  int nj() const { return (has_debye ? 3 : 0) + (has_lorentz ? 6 : 0); }
  int* src2 = nullptr;
  long long* src3 = nullptr;
  float* wave = nullptr;
  long long wave_stride = 0;
  short* probe_head = nullptr;
  short* probe_next = nullptr;
  int* probe_cell = nullptr;   // [P] 2D box-linear cell of each probe, -1 outside
  long long* probe_cell3 = nullptr;  // [P] 3D box-linear cell of each probe, -1 outside
  float* traces = nullptr;
  float* partial = nullptr;
  float2* tw = nullptr;        // DFT twiddles [T][F] for tw_freqs
  float2* S = nullptr;         // DFT output [F][P][B]
  float2* S_host = nullptr;    // its pinned host staging (hipHostMalloc), same size
  std::vector<double> tw_freqs;
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};   // one per starting buffer parity
  int parity = 0;              // ping-pong buffer holding the current state

  // adjoint: forward checkpoints every ck_every steps, e store of one segment,
  // adjoint state (ping-pong), gradient sums
  int ck_every = 0, ck_count = 0;
  std::vector<float*> ck;      // ck_count x 3 fields (ez, hx, hy) [B][plane]
  float* estore = nullptr;     // (ck_every + 1) x [B][plane]
  float* adj[2][3] = {{nullptr}};
  float* g_ca = nullptr;
  float* g_cbi = nullptr;
  float* d_dtr = nullptr;
  float* adj3[6][3] = {{nullptr}};  // 3D: lam, hbar, q, Debye, Lorentz, Lorentz-prev adjoints
  // 3D blocked adjoint: lam then hbar, each component two nbox-float sets
  // (12 nbox), the packed adjoint currents (2 nj nbox) and the gradient sums
  // per TX ([2][B][plane])
  float* adjtb = nullptr;
  float* adjtb_j = nullptr;
  float* g3tb = nullptr;
  bool ck_valid = false;
  int estore_seg = -1;         // segment whose forward E the store holds, or -1
  int estore_off = 0;          // ... from this slot on (3D: estore_x for the tail segment)
  // 3D: extra E-store slots.  The store holds every + 1 + estore_x slots; the
  // checkpointed forward leaves the last segment in slots [0, every] and the
  // top estore_x steps of the segment before it (estore_tail) behind them, at
  // slot estore_x + j for its step j, so that segment's adjoint recomputes only
  // its first every - estore_x steps (into slots estore_x + [0, m], freed by
  // the last segment).  estore_extra (tuning key 25): -1 = as many as the free
  // device memory holds after what the adjoint still allocates, else that many.
  int estore_x = 0;
  int estore_tail = -1;
  int estore_m = 0;            // the tail segment's steps not kept (recomputed by the adjoint)
  int estore_extra = -1;

  // FP64 forward of the gradient path (tuning key 5): fields, maps, waveform,
  // traces and checkpoints in doubles; the adjoint itself stays FP32
  int fwd64 = 0;
  double* f64[2][3] = {{nullptr}};
  double *ca64 = nullptr, *cb64 = nullptr, *wave64 = nullptr, *traces64 = nullptr;
  std::vector<double> wave_h;  // the waveform as given (FP64), [T] or [B][T]
  std::vector<double*> ck64;   // ck_count x 3 FP64 checkpoints
  bool maps64_dirty = true;
  bool traces64_valid = false; // the last run was the FP64 forward

  // 2D non-dispersive path: the FP64 material maps (full grid) on the device;
  // the coefficient maps and the gradient chain rule are formed from them there
  double* eps64 = nullptr;
  double* sig64 = nullptr;
  double* grad64 = nullptr;    // [2][full_cells] dL/d eps_r, dL/d sigma

  // difference (residual) form of the 2D gradient path: a reference medium
  // (the misfit's target), its run's E at every step and its state at every
  // checkpoint, and the maps of the source term E - E_ref is stepped with
  bool has_ref = false;
  bool ref_dirty = true;       // the reference run is missing or stale
  int ref_every = 0;
  double* ref_eps64 = nullptr;
  double* ref_sig64 = nullptr;
  float *ca_ref = nullptr, *cb_ref = nullptr, *d_a = nullptr, *d_b = nullptr;
  unsigned char* diff_tile = nullptr;
  int diff_tile_n = 0;
  // the flagged tiles and the others as lists ([n_diff] then [n_plain] ints):
  // by default one difference launch per step takes the flagged tiles first;
  // key 23 = 2 runs them with the difference kernel on an extra stream while
  // the plain kernel (no source, 80 VGPRs against 127) steps the rest
  int* tile_lists = nullptr;
  int n_diff_tiles = 0, n_plain_tiles = 0;
  int split_diff = 1;          // tuning key 23: 0 = one difference kernel over every tile
  int tx_streams = 2;          // tuning key 27: TX groups of the 2D blocked kernels on 1, 2 or 4 streams
  int tx_streams3d = 1;        // tuning key 28: the same for the 3D blocked forward
  int tx_streams3d_adj = 1;    // tuning key 29: and for the 3D blocked adjoint
  // extra streams (default priority): xs[0] also serves the difference
  // form's split modes; ev_fork / ev_join[i] fork work onto xs[i] and join it
  hipStream_t xs[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[3] = {nullptr, nullptr, nullptr};
  float* e_ref = nullptr;      // (T + 1) slots of [B][plane]: E_ref^s
  std::vector<float*> ref_ck;  // (nck + 1) x 3: reference state every ref_every steps
  std::vector<float> ref_traces;   // [B][P][T]
  bool diff_active = false;    // launch_range steps the difference form

  // piecewise-constant 2D media: per-cell material id and the (ca, cb, k1,
  // k2dt) table the blocked four-wavefront kernels read instead of the maps
  unsigned char* mat_id = nullptr;
  float4* mat_table = nullptr;
  int n_mat = 0;               // 0: no id map (too many distinct cells, or switched off)
  int mat_ids_on = 1;          // tuning key 22: 0 off, 1 2D kernels + 3D adjoint, 2 also the 3D forward

  bool mats_dirty = true;
  bool has_run = false;
  int bb = 4;
  int use_graph = 1;
  int tb_k = 0;                // 2D temporal depth (0: one step per launch)
  int tb_r = 16;               // 2D rows per wavefront window
  int tb_bpw = 1;              // 2D batch planes per wavefront
  int fused3d = 4;             // 3D step form: 0 = H then E kernels in place; 1/2 = one fused H+E launch
                               // per step, 3 = its marching form (ping-pong fields); 4 (default) = the
                               // temporally blocked kernel, tb3d_k steps per launch (the whole state ping-pongs)
  int march_len = 24;          // 3D marching form: x-slabs per workgroup
  int xcd_order = 1;           // 2D blocked kernels: contiguous tile ranges per XCD
  int adj_r = 0;               // 2D blocked adjoint: window rows (0: 32 for the four-wavefront
                               // reversible kernel, 16 for the one-wavefront kernels)
  int adj_bpw = 4;             // 2D blocked adjoint: batch planes per wavefront
  int adj_waves = 4;           // 2D reversible adjoint: wavefronts per window (1, 4 or 8)
  int adj_prefetch = 0;        // 2D reversible adjoint: next TX's planes loaded ahead (key 24)
  // device S-misfit of ceep_fdtd_gradient_misfit: FP64 twiddles [T][F] for
  // misfit_freqs, the misfit D [F][P][B], S_target, and (loss, dL/dtrace bound)
  double2* tw64 = nullptr;
  std::vector<double> misfit_freqs;
  double2* misfit_d = nullptr;
  double2* misfit_tgt = nullptr;
  double* misfit_out = nullptr;
  float* pin_dtr = nullptr;    // pinned host staging: dL/dtrace (FP32) ...
  float* pin_g = nullptr;      // ... and the two gradient planes
  size_t pin_dtr_n = 0, pin_g_n = 0;
  int adj_k = 0;               // 2D blocked adjoint: reversed steps per launch (0: 6 for the
                               // four-wavefront reversible kernel on 32-row windows, 4 otherwise)
  int fwd_waves = 1;           // 2D blocked forward: wavefronts per window (1, 4 or 8)
  int xcd3d = 0;               // 3D H/E kernels: block order (0 dispatch, 1/2 per XCD)
  int tb3d_k = 2;              // 3D blocked form: steps per launch (1 or 2)
  int tx_group3d = 4;          // 3D blocked form, block order 3: TX per group
  int adj3d_k = 2;             // 3D adjoint: reversed steps per blocked launch (0: the A/B/C kernels)
  int adj3d_order = 1;         // 3D blocked adjoint: workgroup order (0 dispatch, 1 per XCD, 2 TX fastest)
  int adj_rev = 1;             // 2D FP32 blocked adjoint rebuilds the forward state backward
  bool ck_rev = false;         // the checkpoints were taken for the reversible adjoint
  float ch = 0.f, idx = 0.f;

  // physical-CPML mode (CEEP_BOX_CPML): profile parameters of the reference's
  // CPML class (absorbing.py:226-230 defaults) and the psi state
  int cpml_order = 4;
  double cpml_sigma_factor = 2.0, cpml_alpha_max = 0.05, cpml_kappa_max = 1.0;
  // physical-CPML psi arrays (field layout): 2D hyx, hxy, ezx, ezy; 3D hxy, hxz,
  // hyz, hyx, hzx, hzy, exy, exz, eyz, eyx, ezx, ezy
  float* psi[12] = {nullptr};
  bool cpml() const { return box_mode == CEEP_BOX_CPML; }

  long long box_floats() const { return (long long)B * plane; }
  // field buffer holding the current state (3D fields stay in buffer 0 with
  // the two-kernel in-place update)
  int fbuf() const { return cpml() ? 0 : (ndim == 2 || form3d()) ? parity : 0; }
  int poles() const { return (has_debye ? 1 : 0) | (has_lorentz ? 2 : 0); }
  // the 3D step form in effect: the blocked form (4) addresses a TX's two
  // state sets and its packed maps with 32-bit byte offsets, so beyond those
  // limits the handle steps with the one-step kernels (form 0) instead
  int form3d() const {
    const bool fits = ((long long)B * plane + plane) * 4 < (1LL << 31) && plane * 36 < (1LL << 31);
    return fused3d == 4 && !fits ? 0 : fused3d;
  }
  // 3D temporally blocked form: every state array holds two sets (fields
  // f[0]/f[1], J packed per cell at set * nj * box_floats()); set `parity` is
  // live, and each set keeps its Lorentz J as (newest, one step older)
  bool tb3d() const { return ndim == 3 && !cpml() && form3d() == 4; }
  // 3D blocked adjoint: runs with the blocked forward (its packed maps) while
  // an E-store slot of three components fits its 32-bit byte offsets
  bool adjtb3d() const {
    return tb3d() && adj3d_k > 0 && (2 * box_floats() + plane) * 4 < (1LL << 31);
  }
  // the temporally blocked 2D kernel covers the plain and Debye cases
  bool use_tb() const {
    return ndim == 2 && tb_k > 0 && !has_lorentz && !cpml() && ceep::tb2d_fits(plane);
  }
  // 2D gradient path: reversible reconstruction instead of recompute + E store
  bool use_rev() const { return adj_rev && use_tb(); }
  int launches_per_run() const {
    return use_tb() ? (T + tb_k - 1) / tb_k : tb3d() ? (T + tb3d_k - 1) / tb3d_k : T;
  }
  long long full_index(int64_t x, int64_t y, int64_t z) const {
    return ndim == 2 ? x * N[1] + y : (x * N[1] + y) * N[2] + z;
  }
  long long box_index(int a, int b, int c) const {
    return ndim == 2 ? (long long)a * pitch + b
                     : (long long)a * row_block + (long long)b * pitch + c;
  }
  void free_adjoint() {
    for (auto& p : ck) if (p) (void)hipFree(p);
    ck.clear();
    for (auto& p : ck64) if (p) (void)hipFree(p);
    ck64.clear();
    ck_count = 0;
    ck_valid = false;
    if (estore) (void)hipFree(estore);
    estore = nullptr;
    estore_seg = -1;
    estore_off = 0;
    estore_x = 0;
    estore_tail = -1;
    estore_m = 0;
    for (auto& row : adj)
      for (auto& p : row) { if (p) (void)hipFree(p); p = nullptr; }
    for (auto& row : adj3)
      for (auto& p : row) { if (p) (void)hipFree(p); p = nullptr; }
    for (float** p : {&adjtb, &adjtb_j, &g3tb}) { if (*p) (void)hipFree(*p); *p = nullptr; }
    if (g_ca) (void)hipFree(g_ca);
    if (g_cbi) (void)hipFree(g_cbi);
    if (d_dtr) (void)hipFree(d_dtr);
    g_ca = g_cbi = d_dtr = nullptr;
    for (double2** p : {&tw64, &misfit_d, &misfit_tgt}) { if (*p) (void)hipFree(*p); *p = nullptr; }
    if (misfit_out) (void)hipFree(misfit_out);
    misfit_out = nullptr;
    misfit_freqs.clear();
    if (pin_dtr) (void)hipHostFree(pin_dtr);
    if (pin_g) (void)hipHostFree(pin_g);
    pin_dtr = pin_g = nullptr;
    pin_dtr_n = pin_g_n = 0;
  }
  void free_reference() {
    for (auto& p : ref_ck) if (p) (void)hipFree(p);
    ref_ck.clear();
    for (float** p : {&e_ref, &ca_ref, &cb_ref, &d_a, &d_b}) { if (*p) (void)hipFree(*p); *p = nullptr; }
    if (diff_tile) (void)hipFree(diff_tile);
    diff_tile = nullptr;
    if (tile_lists) (void)hipFree(tile_lists);
    tile_lists = nullptr;
    n_diff_tiles = n_plain_tiles = 0;
    diff_tile_n = 0;
    ref_traces.clear();
    ref_dirty = true;
    ref_every = 0;
  }
  void drop_graph() {
    for (auto& g : gexec) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }
};

namespace {

int upload_f32(float* dst, const std::vector<float>& v, hipStream_t s) {
  CEEP_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice, s));
  return CEEP_OK;
}

// FP64 coefficient maps (fdtd_2d_batched.py:152-165, fdtd_3d_batched.py:282-296,
// dispersive.py:253-299, core/grid.py:178-184) over the box, stored as FP32.
// Material ids of piecewise-constant media: the distinct map records of the
// box, id 0 the all-zero record of cells off the E interior; none when there
// are more than the kernels' table holds, or the feature is off.  2D (host-
// formed maps, no Lorentz pole): records (ca, cb, k1, k2dt), one float4 per
// material.  3D blocked form without mu_r / sigma_m maps: the packed record
// (ca, cb, Debye k1, k2dt, Lorentz k1, k2dt, k3; absent poles take no room)
// padded to kMat3Floats.  The tables hold the exact FP32 map values, so
// results are bit-identical to the maps'.
int build_material_ids(ceep_fdtd* h, const std::vector<float>& ca, const std::vector<float>& cb,
                       const std::vector<float>& k1, const std::vector<float>& k2,
                       const std::vector<float>& l1, const std::vector<float>& l2,
                       const std::vector<float>& l3, bool magnetic) {
  dfree(h->mat_id);
  dfree(h->mat_table);
  h->n_mat = 0;
  if (!h->mat_ids_on || ca.empty() || h->cpml() || h->plane >= (1LL << 29)) return CEEP_OK;
  const bool d2 = h->ndim == 2 && !h->has_lorentz;
  const bool d3 = h->tb3d() && !magnetic;
  if (!d2 && !d3) return CEEP_OK;
  const int stride = d2 ? 4 : ceep::kMat3Floats;
  const int max_mat = d2 ? ceep::kMaxMaterials : ceep::kMaxMat3;
  using Key = std::array<uint32_t, 8>;
  struct Hash {
    size_t operator()(const Key& k) const {
      uint64_t x = 1469598103934665603ULL;
      for (uint32_t u : k) x = (x ^ u) * 1099511628211ULL;
      return (size_t)x;
    }
  };
  auto bits = [](float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; };
  std::unordered_map<Key, int, Hash> ids;
  std::vector<float> table((size_t)stride, 0.f);
  ids.emplace(Key{}, 0);
  std::vector<unsigned char> id((size_t)h->plane, 0);
  std::vector<const std::vector<float>*> cols = {&ca, &cb};
  for (const auto* c : {&k1, &k2, &l1, &l2, &l3})
    if (!c->empty()) cols.push_back(c);
  float rec[8];
  for (long long o = 0; o < h->plane; ++o) {
    Key k{};
    for (size_t j = 0; j < cols.size(); ++j) {
      rec[j] = (*cols[j])[o];
      k[j] = bits(rec[j]);
    }
    auto it = ids.find(k);
    if (it == ids.end()) {
      if ((int)ids.size() == max_mat) return CEEP_OK;   // not piecewise constant
      it = ids.emplace(k, (int)ids.size()).first;
      table.resize(table.size() + stride, 0.f);
      std::memcpy(&table[table.size() - stride], rec, cols.size() * sizeof(float));
    }
    id[(size_t)o] = (unsigned char)it->second;
  }
  int rc;
  if ((rc = dalloc(&h->mat_id, id.size())) || (rc = dalloc(&h->mat_table, table.size() / 4)))
    return rc;
  CEEP_HIP(hipMemcpyAsync(h->mat_id, id.data(), id.size(), hipMemcpyHostToDevice, h->stream));
  CEEP_HIP(hipMemcpyAsync(h->mat_table, table.data(), table.size() * sizeof(float),
                          hipMemcpyHostToDevice, h->stream));
  CEEP_HIP(hipStreamSynchronize(h->stream));
  h->n_mat = (int)ids.size();
  return CEEP_OK;
}

// 2D without poles: ca / cb (and, with a reference medium, its maps and the
// difference-form maps) formed on the device from the FP64 maps
int commit_materials_2d_device(ceep_fdtd* h) {
  hipStream_t s = h->stream;
  int rc;
  const size_t nf = (size_t)h->full_cells;
  if (!h->eps64 && (rc = dalloc(&h->eps64, nf))) return rc;
  if (!h->sig64 && (rc = dalloc(&h->sig64, nf))) return rc;
  CEEP_HIP(hipMemcpyAsync(h->eps64, h->eps_r.data(), nf * sizeof(double), hipMemcpyHostToDevice, s));
  CEEP_HIP(hipMemcpyAsync(h->sig64, h->sigma_e.data(), nf * sizeof(double), hipMemcpyHostToDevice, s));
  ceep::Coef2DArgs a{};
  a.eps_r = h->eps64; a.sigma = h->sig64;
  a.ny_full = h->N[1]; a.org = h->org;
  a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch;
  a.cpml_interior = h->cpml() ? 1 : 0;
  a.dt = h->dt;
  a.ca = h->ca; a.cb = h->cb;
  if (h->has_ref) {
    for (float** p : {&h->ca_ref, &h->cb_ref, &h->d_a, &h->d_b})
      if (!*p && (rc = dalloc(p, (size_t)h->plane))) return rc;
    a.eps_ref = h->ref_eps64; a.sigma_ref = h->ref_sig64;
    a.ca_ref = h->ca_ref; a.cb_ref = h->cb_ref; a.d_a = h->d_a; a.d_b = h->d_b;
  }
  CEEP_HIP(ceep::launch_coef2d(a, s));
  return CEEP_OK;
}

int commit_materials(ceep_fdtd* h) {
  if (!h->mats_dirty) return CEEP_OK;
  const long long nbox = h->plane;
  const bool dev2d = h->ndim == 2 && !h->has_debye && !h->has_lorentz;
  std::vector<float> ca(dev2d ? 0 : nbox, 0.f), cb(dev2d ? 0 : nbox, 0.f);
  std::vector<float> da, dbi, dk1, dk2, lk1, lk2, lk3;
  if (h->ndim == 3 && h->magnetic) { da.assign(nbox, 1.f); dbi.assign(nbox, 0.f); }
  if (h->has_debye) { dk1.assign(nbox, 0.f); dk2.assign(nbox, 0.f); }
  if (h->has_lorentz) { lk1.assign(nbox, 0.f); lk2.assign(nbox, 0.f); lk3.assign(nbox, 0.f); }
  const double dt = h->dt, inv = 1.0 / h->dx;
  const int mz = h->ndim == 3 ? h->M[2] : 1;
  for (int a = 0; a < (dev2d ? 0 : h->M[0]); ++a)
    for (int b = 0; b < h->M[1]; ++b)
      for (int c = 0; c < mz; ++c) {
        const long long g = h->full_index(a + h->org, b + h->org, h->ndim == 3 ? c + h->org : 0);
        const long long o = h->ndim == 2 ? h->box_index(a, b, 0) : h->box_index(a, b, c);
        double eps_add = 0.0;
        if (h->has_debye) {
          const double tau = h->deb_tau[g];
          if (tau > 0) {
            const double denom = 2.0 * tau + dt;
            const double k1 = (2.0 * tau - dt) / denom;
            const double k2 = (2.0 * kEps0 * h->deb_de[g] * dt) / denom;
            eps_add += k2 / 2.0;
            dk1[o] = (float)k1;
            dk2[o] = (float)(k2 / dt);
          }
        }
        if (h->has_lorentz) {
          const double w0 = h->lor_w0[g];
          if (w0 > 0) {
            const double d = h->lor_dl[g], w02 = w0 * w0, dt2 = dt * dt;
            const double denom = 4.0 + 2.0 * d * dt + w02 * dt2;
            const double k1 = (8.0 - 2.0 * w02 * dt2) / denom;
            const double k3 = -(4.0 - 2.0 * d * dt + w02 * dt2) / denom;
            const double k2 = 2.0 * kEps0 * h->lor_de[g] * w02 * dt2 / denom;
            eps_add += k2 / 2.0;
            lk1[o] = (float)k1;
            lk3[o] = (float)k3;
            lk2[o] = (float)(k2 / dt);
          }
        }
        double eps = h->eps_r[g] * kEps0;
        if (h->has_debye || h->has_lorentz) eps = eps + eps_add;
        const double s = h->sigma_e[g] * dt / (2.0 * eps);
        // E is stepped on the interior only; elsewhere every E-side
        // coefficient is 0 so the kernels' uniform update keeps E at 0
        const bool interior =
            h->cpml() ? (a >= 1 && b >= 1 && (h->ndim == 2 || c >= 1))   // [1:, 1:(, 1:)]
                                                  // (fdtd_2d.py:168, fdtd_3d.py:154)
                      : (a >= 1 && a <= h->M[0] - 2 && b >= 1 && b <= h->M[1] - 2 &&
                         (h->ndim == 2 || (c >= 1 && c <= mz - 2)));
        ca[o] = interior ? (float)((1.0 - s) / (1.0 + s)) : 0.f;
        cb[o] = interior ? (float)((dt / eps) / (1.0 + s)) : 0.f;
        if (!interior) {
          if (h->has_debye) { dk1[o] = 0.f; dk2[o] = 0.f; }
          if (h->has_lorentz) { lk1[o] = 0.f; lk2[o] = 0.f; lk3[o] = 0.f; }
        }
        if (!da.empty()) {
          const double mu = h->mu_r[g] * kMu0;
          const double sm = h->sigma_m[g] * dt / (2.0 * mu);
          da[o] = (float)((1.0 - sm) / (1.0 + sm));
          dbi[o] = (float)((dt / mu) / (1.0 + sm) * inv);
        }
      }
  hipStream_t s = h->stream;
  int rc;
  if (dev2d) {
    if ((rc = commit_materials_2d_device(h))) return rc;
  } else if ((rc = upload_f32(h->ca, ca, s)) || (rc = upload_f32(h->cb, cb, s))) {
    return rc;
  }
  if ((rc = build_material_ids(h, ca, cb, dk1, dk2, lk1, lk2, lk3, !da.empty()))) return rc;
  dfree(h->da); dfree(h->dbi); dfree(h->dk1); dfree(h->dk2);
  dfree(h->lk1); dfree(h->lk2); dfree(h->lk3);
  if (!da.empty()) {
    if ((rc = dalloc(&h->da, nbox)) || (rc = dalloc(&h->dbi, nbox))) return rc;
    if ((rc = upload_f32(h->da, da, s)) || (rc = upload_f32(h->dbi, dbi, s))) return rc;
  }
  if (h->has_debye) {
    if ((rc = dalloc(&h->dk1, nbox)) || (rc = dalloc(&h->dk2, nbox))) return rc;
    if ((rc = upload_f32(h->dk1, dk1, s)) || (rc = upload_f32(h->dk2, dk2, s))) return rc;
  }
  if (h->has_lorentz) {
    if ((rc = dalloc(&h->lk1, nbox)) || (rc = dalloc(&h->lk2, nbox)) ||
        (rc = dalloc(&h->lk3, nbox)))
      return rc;
    if ((rc = upload_f32(h->lk1, lk1, s)) || (rc = upload_f32(h->lk2, lk2, s)) ||
        (rc = upload_f32(h->lk3, lk3, s)))
      return rc;
  }
  // 3D blocked form: packed maps, and the ADE currents packed per cell in two
  // sets (buffers of an unchanged size are kept)
  const size_t want_mp = h->tb3d() ? (size_t)nbox * ceep::map_pack(h->poles(), !da.empty()).n : 0;
  const size_t want_jp = h->tb3d() ? 2 * (size_t)h->nj() * h->box_floats() : 0;
  if (want_mp != h->mpack_n) { dfree(h->mpack); h->mpack_n = 0; }
  if (want_jp != h->jpack_n) { dfree(h->jpack); h->jpack_n = 0; }
  if (h->tb3d()) {
    // only the groups the kernels read for this material set (no room for
    // absent poles or, without mu_r / sigma_m maps, for da and dbi)
    const ceep::MapPack mk = ceep::map_pack(h->poles(), !da.empty());
    std::vector<float> mp((size_t)nbox * mk.n, 0.f);
    for (long long o = 0; o < nbox; ++o) {
      float* q = &mp[(size_t)o * mk.n];
      q[0] = ca[o]; q[1] = cb[o];
      if (h->has_debye) { q[mk.dk] = dk1[o]; q[mk.dk + 1] = dk2[o]; }
      if (h->has_lorentz) { q[mk.lk] = lk1[o]; q[mk.lk + 1] = lk2[o]; q[mk.lk + 2] = lk3[o]; }
      if (!da.empty()) { q[mk.mag] = da[o]; q[mk.mag + 1] = dbi[o]; }
    }
    if (!h->mpack && (rc = dalloc(&h->mpack, mp.size()))) return rc;
    h->mpack_n = mp.size();
    if ((rc = upload_f32(h->mpack, mp, s))) return rc;
    CEEP_HIP(hipStreamSynchronize(s));   // mp goes out of scope
    if (want_jp) {
      if (!h->jpack && (rc = dalloc(&h->jpack, want_jp))) return rc;
      h->jpack_n = want_jp;
      CEEP_HIP(hipMemsetAsync(h->jpack, 0, want_jp * sizeof(float), s));
    }
  }
  // ADE current buffers follow the pole set (the blocked 3D form uses jpack)
  const int ncomp = h->tb3d() ? 0 : h->ndim == 2 ? 1 : 3;
  for (int k = 0; k < 3; ++k) {
    dfree(h->jd[k]);
    dfree(h->jl[0][k]);
    dfree(h->jl[1][k]);
  }
  const size_t jbytes = (size_t)h->box_floats() * sizeof(float);
  for (int k = 0; k < ncomp; ++k) {
    if (h->has_debye) {
      if ((rc = dalloc(&h->jd[k], h->box_floats()))) return rc;
      CEEP_HIP(hipMemsetAsync(h->jd[k], 0, jbytes, s));
    }
    if (h->has_debye && h->ndim == 2) {   // ping-pong partner for the blocked kernel
      if ((rc = dalloc(&h->jd[1], h->box_floats()))) return rc;
      CEEP_HIP(hipMemsetAsync(h->jd[1], 0, jbytes, s));
    }
    if (h->has_lorentz) {
      for (int i = 0; i < 2; ++i) {
        if ((rc = dalloc(&h->jl[i][k], h->box_floats()))) return rc;
        CEEP_HIP(hipMemsetAsync(h->jl[i][k], 0, jbytes, s));
      }
    }
  }
  CEEP_HIP(hipStreamSynchronize(s));
  h->drop_graph();
  h->mats_dirty = false;
  return CEEP_OK;
}

int zero_state(ceep_fdtd* h) {
  hipStream_t s = h->stream;
  const size_t bytes = (size_t)h->box_floats() * sizeof(float);
  const int nf = h->ndim == 2 ? 3 : 6;
  for (int k = 0; k < nf; ++k)   // f[1][k] = f[0][k] + box_floats(): one clear per pair
    CEEP_HIP(hipMemsetAsync(h->f[0][k], 0, 2 * bytes, s));
  for (int k = 0; k < 12; ++k)
    if (h->psi[k]) CEEP_HIP(hipMemsetAsync(h->psi[k], 0, bytes, s));
  for (int k = 0; k < 3; ++k) {
    if (h->jd[k]) CEEP_HIP(hipMemsetAsync(h->jd[k], 0, bytes, s));
    if (h->jl[0][k]) CEEP_HIP(hipMemsetAsync(h->jl[0][k], 0, bytes, s));
    if (h->jl[1][k]) CEEP_HIP(hipMemsetAsync(h->jl[1][k], 0, bytes, s));
  }
  if (h->jpack) CEEP_HIP(hipMemsetAsync(h->jpack, 0, 2 * h->nj() * bytes, s));
  h->parity = 0;
  return CEEP_OK;
}

int launch_range(ceep_fdtd* h, int start, int step0, int nsteps, float* e_save);

// The handle's second stream (default priority) and the two events that fork
// work onto it and join it back, created once.
int ensure_streams(ceep_fdtd* h, int n) {
  for (int i = 0; i < n && i < 3; ++i) {
    if (!h->xs[i]) {
      hipError_t e = hipStreamCreateWithFlags(&h->xs[i], hipStreamNonBlocking);
      if (e != hipSuccess) { h->xs[i] = nullptr; return fail(CEEP_EHIP, "extra stream"); }
    }
    if (!h->ev_join[i]) CEEP_HIP(hipEventCreateWithFlags(&h->ev_join[i], hipEventDisableTiming));
  }
  if (!h->ev_fork) CEEP_HIP(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
  return CEEP_OK;
}

// Work queued on s so far happens before anything queued on the first n extra streams next.
int fork_streams(ceep_fdtd* h, hipStream_t s, int n) {
  int rc = ensure_streams(h, n);
  if (rc) return rc;
  CEEP_HIP(hipEventRecord(h->ev_fork, s));
  for (int i = 0; i < n; ++i) CEEP_HIP(hipStreamWaitEvent(h->xs[i], h->ev_fork, 0));
  return CEEP_OK;
}

// Work queued on the first n extra streams happens before anything queued on s next.
int join_streams(ceep_fdtd* h, hipStream_t s, int n) {
  for (int i = 0; i < n; ++i) {
    CEEP_HIP(hipEventRecord(h->ev_join[i], h->xs[i]));
    CEEP_HIP(hipStreamWaitEvent(s, h->ev_join[i], 0));
  }
  return CEEP_OK;
}

// TX groups of the blocked kernels: g of G takes whole chunks of bpw planes,
// chunks [c0, c0 + nc), planes [b0, b0 + nb); empty groups are dropped.
struct TxGroup { int b0, nb, c0, nc; };
std::vector<TxGroup> tx_groups(int B, int bpw, int G) {
  const int chunks = (B + bpw - 1) / bpw;
  std::vector<TxGroup> out;
  for (int g = 0; g < G; ++g) {
    const int c0 = (int)((long long)chunks * g / G), c1 = (int)((long long)chunks * (g + 1) / G);
    if (c1 <= c0) continue;
    const int b0 = c0 * bpw, b1 = std::min(B, c1 * bpw);
    out.push_back(TxGroup{b0, b1 - b0, c0, c1 - c0});
  }
  return out;
}

// The buffer set (ping-pong parity) live after launch_range(h, start, ., nsteps):
// the 3D blocked form flips it once per launch (tb3d_k steps), the other forms once per step.
int range_end(const ceep_fdtd* h, int start, int nsteps) {
  return start ^ ((h->tb3d() ? (nsteps + h->tb3d_k - 1) / h->tb3d_k : nsteps) & 1);
}

// CPML profiles of CPML._compute_pml_params (absorbing.py:266-311) in FP64,
// rounded once: b, c at the E (integer) and H (half-integer) depths of a slab
// of thickness n cells (dx == dy, so one table serves both axes).
void cpml_profiles(const ceep_fdtd* h, float* be, float* ce, float* bh, float* ch) {
  const int n = h->n;
  const double dt = h->dt, d = h->dx, ord = (double)h->cpml_order;
  const double sigma_max = h->cpml_sigma_factor * (h->cpml_order + 1) / (150.0 * M_PI * d);
  for (int k = 0; k < n; ++k) {
    for (int half = 0; half < 2; ++half) {
      const double rho = half ? (k + 0.5) / n : (double)k / n;
      const double sg = sigma_max * std::pow(rho, ord);
      const double kap = 1.0 + (h->cpml_kappa_max - 1.0) * std::pow(rho, ord);
      const double al = h->cpml_alpha_max * (1.0 - rho);
      const double b = std::exp(-(sg / kap + al) * dt / kEps0);
      const double c = std::fabs(sg) < 1e-20 ? 0.0 : sg / (sg * kap + kap * kap * al) * (b - 1.0);
      if (half) { bh[k] = (float)b; ch[k] = (float)c; }
      else { be[k] = (float)b; ce[k] = (float)c; }
    }
  }
}

// Enqueue total_steps steps starting from ping-pong buffer `start`.
int launch_steps(ceep_fdtd* h, int start) {
  return launch_range(h, start, 0, h->T, nullptr);
}

// Enqueue steps [step0, step0 + nsteps) from ping-pong buffer `start`.  With
// e_save (2D), E after step s is written to slot s + 1 - step0 of the
// [slot][B][plane] store (by the blocked kernel on its output windows only,
// i.e. on the stepped box).
int launch_range(ceep_fdtd* h, int start, int step0, int nsteps, float* e_save) {
  const int step_end = step0 + nsteps;
  hipStream_t s = h->stream;
  const int poles = h->poles();
  if (h->use_tb()) {
    ceep::TB2DArgs a{};
    a.xcd_order = h->xcd_order;
    a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch; a.B = h->B; a.bstride = h->plane;
    a.ca = h->ca; a.cb = h->cb; a.d_k1 = h->dk1; a.d_k2dt = h->dk2;
    if (h->n_mat > 0) { a.mat_id = h->mat_id; a.mat_table = h->mat_table; a.n_mat = h->n_mat; }
    a.ch = h->ch; a.idx = h->idx;
    a.src_cell = h->src2; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
    a.T = h->T; a.P = h->P; a.probe_cell = h->probe_cell; a.traces = h->traces;
    const int K = h->tb_k, R = h->tb_r;
    const bool mw = h->fwd_waves == 4 || h->fwd_waves == 8;
    const bool shape_ok = mw ? ceep::step2d_mw_shape_ok(R, K, h->fwd_waves)
                             : (R == 16 || R == 24 || R == 32) && K >= 1 && K <= 4;
    if (!shape_ok)
      return fail(CEEP_EINVAL, "2D blocked launch shape: one wavefront per window takes 16, 24 or 32 "
                               "rows and depth 1-4; four take 32, 48 or 64 rows and depth 4, 6 or 8; "
                               "eight take 32 rows at depth 4, 48 at 4 or 6, 64 at 4, 6 or 8");
    a.tiles_x = (h->M[0] + (R - 2 * K) - 1) / (R - 2 * K);
    a.tiles_y = (h->M[1] + (64 - 2 * K) - 1) / (64 - 2 * K);
    a.bpw = h->tb_bpw;
    a.ntasks = a.tiles_x * a.tiles_y * ((h->B + a.bpw - 1) / a.bpw);
    int kpoles = h->poles();
    if (h->diff_active) {   // E - E_ref with the reference run's source term
      if (!mw || kpoles != 0 || !h->e_ref || !h->diff_tile ||
          h->diff_tile_n != a.tiles_x * a.tiles_y)
        return fail(CEEP_ESTATE, "difference form: needs the four-wavefront blocked forward, "
                                 "non-dispersive media and a reference run for this launch shape");
      kpoles = ceep::kPolesDiff;
      a.d_a = h->d_a; a.d_b = h->d_b; a.e_ref = h->e_ref; a.ref_slot = h->box_floats();
      a.diff_tile = h->diff_tile;
    }
    // difference form (key 23): 1 = one difference launch over the tiles
    // with the flagged ones first (they take longest and start earliest),
    // 2 = the flagged tiles with the difference kernel on an extra stream and the
    // others with the plain kernel (no source) on stream, 3 = the same two
    // launches one after the other on stream
    const bool lists = kpoles == ceep::kPolesDiff && h->tile_lists && !e_save &&
                       h->n_diff_tiles > 0 && h->n_plain_tiles > 0;
    const bool split = lists && (h->split_diff == 2 || h->split_diff == 3);
    const bool two_streams = split && h->split_diff == 2;
    ceep::TB2DArgs pa = a;
    const int chunks = (h->B + a.bpw - 1) / a.bpw;
    if (lists && h->split_diff == 1) {
      a.tile_list = h->tile_lists;
      a.n_list = h->n_diff_tiles + h->n_plain_tiles;
    }
    if (split) {
      a.tile_list = h->tile_lists;
      a.n_list = h->n_diff_tiles;
      a.ntasks = h->n_diff_tiles * chunks;
      pa.tile_list = h->tile_lists + h->n_diff_tiles;
      pa.n_list = h->n_plain_tiles;
      pa.ntasks = h->n_plain_tiles * chunks;
      pa.no_inject = 1;
      pa.mat_id = nullptr;   // the difference form steps with the full maps
      pa.d_a = pa.d_b = nullptr;
      pa.e_ref = nullptr;
      pa.diff_tile = nullptr;
      // the extra stream starts after everything already queued on stream
      if (two_streams) {
        int rc2 = fork_streams(h, s, 1);
        if (rc2) return rc2;
      }
    }
    // TX groups on separate streams (key 27): the groups share no state, so
    // the extra streams join once before the range and rejoin after it, and
    // each stream's launches fill the partly idle last rounds of the others'
    const std::vector<TxGroup> groups =
        (h->tx_streams > 1 && !split) ? tx_groups(h->B, a.bpw, h->tx_streams) : std::vector<TxGroup>();
    if (groups.size() > 1) {
      const int nx = (int)groups.size() - 1;
      int rc2 = fork_streams(h, s, nx);
      if (rc2) return rc2;
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const TxGroup& g = groups[gi];
        const long long po = (long long)g.b0 * h->plane;
        auto at = [po](float* q) { return q ? q + po : q; };
        ceep::TB2DArgs ha = a;
        ha.B = g.nb;
        ha.ntasks = (a.tile_list ? a.n_list : a.tiles_x * a.tiles_y) * g.nc;
        ha.src_cell = a.src_cell + g.b0;
        ha.wave = a.wave + (long long)g.b0 * a.wave_stride_b;
        ha.traces = a.traces + (long long)g.b0 * h->P * h->T;
        if (ha.e_ref) ha.e_ref = a.e_ref + po;
        const hipStream_t hs = gi ? h->xs[gi - 1] : s;
        int cur = start;
        for (int st = step0; st < step_end; st += K) {
          const int o = cur ^ 1;
          ha.ez_in = at(h->f[cur][0]); ha.hx_in = at(h->f[cur][1]); ha.hy_in = at(h->f[cur][2]);
          ha.ez_out = at(h->f[o][0]); ha.hx_out = at(h->f[o][1]); ha.hy_out = at(h->f[o][2]);
          ha.jd_in = at(h->jd[cur]); ha.jd_out = at(h->jd[o]);
          ha.step0 = st;
          ha.ksteps = std::min(K, step_end - st);
          ha.slot = h->box_floats();
          ha.e_save = e_save ? e_save + (size_t)(st + 1 - step0) * h->box_floats() + po : nullptr;
          if (mw)
            CEEP_HIP(ceep::launch_step2d_mw(ha, kpoles, R, K, h->fwd_waves, hs));
          else
            CEEP_HIP(ceep::launch_step2d_tb(ha, kpoles, R, K, hs));
          cur = o;
        }
      }
      return join_streams(h, s, nx);
    }
    int cur = start;
    for (int st = step0; st < step_end; st += K) {
      const int o = cur ^ 1;
      a.ez_in = h->f[cur][0]; a.hx_in = h->f[cur][1]; a.hy_in = h->f[cur][2];
      a.ez_out = h->f[o][0]; a.hx_out = h->f[o][1]; a.hy_out = h->f[o][2];
      a.jd_in = h->jd[cur]; a.jd_out = h->jd[o];
      a.step0 = st;
      a.ksteps = std::min(K, step_end - st);
      a.slot = h->box_floats();
      a.e_save = e_save ? e_save + (size_t)(st + 1 - step0) * h->box_floats() : nullptr;
      if (split) {
        // both launches read set cur and write set o; each waits for the
        // other's launch of the step before (which read o and wrote cur)
        pa.ez_in = a.ez_in; pa.hx_in = a.hx_in; pa.hy_in = a.hy_in;
        pa.ez_out = a.ez_out; pa.hx_out = a.hx_out; pa.hy_out = a.hy_out;
        pa.jd_in = a.jd_in; pa.jd_out = a.jd_out;
        pa.step0 = a.step0; pa.ksteps = a.ksteps; pa.slot = a.slot; pa.e_save = nullptr;
        CEEP_HIP(ceep::launch_step2d_mw(a, kpoles, R, K, h->fwd_waves, two_streams ? h->xs[0] : s));
        CEEP_HIP(ceep::launch_step2d_mw(pa, 0, R, K, h->fwd_waves, s));
        if (two_streams) {
          CEEP_HIP(hipEventRecord(h->ev_join[0], h->xs[0]));
          CEEP_HIP(hipEventRecord(h->ev_fork, s));
          CEEP_HIP(hipStreamWaitEvent(s, h->ev_join[0], 0));
          CEEP_HIP(hipStreamWaitEvent(h->xs[0], h->ev_fork, 0));
        }
      } else if (mw) {
        CEEP_HIP(ceep::launch_step2d_mw(a, kpoles, R, K, h->fwd_waves, s));
      } else {
        CEEP_HIP(ceep::launch_step2d_tb(a, kpoles, R, K, s));
      }
      cur = o;
    }
  } else if (h->cpml() && h->ndim == 2) {
    ceep::Cpml2DArgs a{};
    a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch; a.B = h->B; a.t = h->n;
    a.bstride = h->plane;
    a.ez = h->f[0][0]; a.hx = h->f[0][1]; a.hy = h->f[0][2];
    a.jd = h->jd[0];
    for (int k = 0; k < 4; ++k) (&a.psi_hyx)[k] = h->psi[k];
    a.ca = h->ca; a.cb = h->cb;
    a.d_k1 = h->dk1; a.d_k2dt = h->dk2; a.l_k1 = h->lk1; a.l_k2dt = h->lk2; a.l_k3 = h->lk3;
    a.dbdy = (float)((h->dt / kMu0) / h->dx);
    a.idx = h->idx;
    a.dtmu = (float)(h->dt / kMu0);
    a.dteps = (float)(h->dt / kEps0);
    cpml_profiles(h, a.be, a.ce, a.bh, a.chh);
    a.src_cell = h->src2; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
    a.T = h->T; a.P = h->P;
    a.probe_head = h->probe_head; a.probe_next = h->probe_next; a.traces = h->traces;
    for (int st = step0; st < step_end; ++st) {
      const int i = (start + st - step0) & 1, o = i ^ 1;
      a.jl_cur = h->jl[i][0]; a.jl_prev = h->jl[o][0];
      a.step = st;
      CEEP_HIP(ceep::launch_cpml2d(a, poles, s));
    }
  } else if (h->ndim == 2) {
    ceep::Step2DArgs a{};
    a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch; a.B = h->B; a.bstride = h->plane;
    a.ca = h->ca; a.cb = h->cb;
    a.d_k1 = h->dk1; a.d_k2dt = h->dk2;
    a.l_k1 = h->lk1; a.l_k2dt = h->lk2; a.l_k3 = h->lk3;
    a.ch = h->ch; a.idx = h->idx;
    a.src_cell = h->src2; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
    a.T = h->T; a.P = h->P;
    a.probe_head = h->probe_head; a.probe_next = h->probe_next; a.traces = h->traces;
    for (int st = step0; st < step_end; ++st) {
      const int i = (start + st - step0) & 1, o = i ^ 1;
      a.ez_in = h->f[i][0]; a.hx_in = h->f[i][1]; a.hy_in = h->f[i][2];
      a.ez_out = h->f[o][0]; a.hx_out = h->f[o][1]; a.hy_out = h->f[o][2];
      a.jl_cur = h->jl[i][0]; a.jl_prev = h->jl[o][0];
      a.jd = h->jd[start];       // updated in place; lives where the run started
      a.step = st;
      a.e_save = e_save ? e_save + (size_t)(st + 1 - step0) * h->box_floats() : nullptr;
      CEEP_HIP(ceep::launch_step2d(a, poles, h->bb, s));
    }
  } else if (h->cpml()) {
    ceep::Cpml3DArgs a{};
    a.Mx = h->M[0]; a.My = h->M[1]; a.Mz = h->M[2]; a.pitch = h->pitch; a.B = h->B; a.t = h->n;
    a.sx = h->row_block; a.bstride = h->plane;
    a.ex = h->f[0][0]; a.ey = h->f[0][1]; a.ez = h->f[0][2];
    a.hx = h->f[0][3]; a.hy = h->f[0][4]; a.hz = h->f[0][5];
    for (int k = 0; k < 3; ++k) a.jd[k] = h->jd[k];
    float** ps = &a.psi_hxy;
    for (int k = 0; k < 12; ++k) ps[k] = h->psi[k];
    a.ca = h->ca; a.cb = h->cb; a.da = h->da; a.dbi = h->dbi;
    a.d_k1 = h->dk1; a.d_k2dt = h->dk2; a.l_k1 = h->lk1; a.l_k2dt = h->lk2; a.l_k3 = h->lk3;
    a.ch = h->ch; a.idx = h->idx;
    a.dtmu = (float)(h->dt / kMu0);
    a.dteps = (float)(h->dt / kEps0);
    cpml_profiles(h, a.be, a.ce, a.bh, a.chh);
    a.src_cell = h->src3; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
    a.T = h->T; a.P = h->P;
    a.probe_head = h->probe_head; a.probe_next = h->probe_next; a.traces = h->traces;
    for (int st = step0; st < step_end; ++st) {
      const int i = (start + st - step0) & 1, o = i ^ 1;
      for (int k = 0; k < 3; ++k) { a.jl_cur[k] = h->jl[i][k]; a.jl_prev[k] = h->jl[o][k]; }
      a.step = st;
      CEEP_HIP(ceep::launch_cpml3d(a, poles, s));
    }
  } else if (h->tb3d()) {
    // two steps per launch (one for an odd remainder), set `cur` -> the other
    ceep::TB3DArgs t{};
    t.Mx = h->M[0]; t.My = h->M[1]; t.Mz = h->M[2]; t.pitch = h->pitch; t.B = h->B;
    t.sx = h->row_block; t.bstride = h->plane; t.nbox = h->box_floats();
    for (int k = 0; k < 3; ++k) { t.e[k] = h->f[0][k]; t.h[k] = h->f[0][3 + k]; }
    t.jp = h->jpack;
    t.mp = h->mpack;
    t.magnetic = h->da != nullptr;
    if (h->n_mat > 0 && h->mat_ids_on == 2) {   // measured slower than the packed maps: opt-in
      t.mat_id = h->mat_id; t.mat_table = reinterpret_cast<const float*>(h->mat_table);
      t.n_mat = h->n_mat;
    }
    t.ch = h->ch; t.idx = h->idx;
    t.src_cell = h->src3; t.wave = h->wave; t.wave_stride_b = h->wave_stride;
    t.T = h->T; t.P = h->P;
    t.probe_head = h->probe_head; t.probe_next = h->probe_next; t.traces = h->traces;
    t.order = h->xcd3d;
    // TX groups on separate streams (key 28), as in 2D: a launch runs 2.8
    // rounds of one-per-CU workgroups, and the other group's launch may take
    // the CUs its last partial round leaves idle (measured slower: the
    // groups' windows then share each XCD's L2)
    std::vector<TxGroup> groups = tx_groups(h->B, 1, std::max(1, h->tx_streams3d));
    const int nx = (int)groups.size() - 1;
    if (nx > 0) {
      int rc2 = fork_streams(h, s, nx);
      if (rc2) return rc2;
    }
    for (int gi = 0; gi <= nx; ++gi) {
      t.b0 = groups[gi].b0;
      t.nb = groups[gi].nb;
      t.tx_group = std::min(h->tx_group3d, t.nb);
      const hipStream_t gs = gi ? h->xs[gi - 1] : s;
      int cur = start;
      for (int st = step0; st < step_end;) {
        const int K = std::min(h->tb3d_k, step_end - st);
        t.in_set = cur;
        t.step = st;
        t.tiles_y = (t.My - 1 + (16 - 2 * K) - 1) / (16 - 2 * K);
        t.tiles_z = (t.Mz - 1 + (64 - 2 * K) - 1) / (64 - 2 * K);
        t.e_save = e_save ? e_save + (size_t)(st + 1 - step0) * 3 * h->box_floats() : nullptr;
        CEEP_HIP(ceep::launch_step3d_tb(t, K, poles, gs));
        cur ^= 1;
        st += K;
      }
    }
    if (nx > 0) return join_streams(h, s, nx);
  } else {
    ceep::Step3DArgs a{};
    a.Mx = h->M[0]; a.My = h->M[1]; a.Mz = h->M[2]; a.pitch = h->pitch; a.B = h->B;
    a.sx = h->row_block; a.bstride = h->plane;
    a.ex = h->f[0][0]; a.ey = h->f[0][1]; a.ez = h->f[0][2];
    a.hx = h->f[0][3]; a.hy = h->f[0][4]; a.hz = h->f[0][5];
    for (int k = 0; k < 3; ++k) a.jd[k] = h->jd[k];
    a.ca = h->ca; a.cb = h->cb; a.da = h->da; a.dbi = h->dbi;
    a.d_k1 = h->dk1; a.d_k2dt = h->dk2;
    a.l_k1 = h->lk1; a.l_k2dt = h->lk2; a.l_k3 = h->lk3;
    a.ch = h->ch; a.idx = h->idx;
    a.src_cell = h->src3; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
    a.T = h->T; a.P = h->P;
    a.probe_head = h->probe_head; a.probe_next = h->probe_next; a.traces = h->traces;
    const int bb = h->bb > 4 ? 4 : h->bb;
    a.e_cstride = h->box_floats();
    a.xcd_order = h->xcd3d;
    for (int st = step0; st < step_end; ++st) {
      const int i = (start + st - step0) & 1, o = i ^ 1;
      for (int k = 0; k < 3; ++k) { a.jl_cur[k] = h->jl[i][k]; a.jl_prev[k] = h->jl[o][k]; }
      a.step = st;
      a.e_save = e_save ? e_save + (size_t)(st + 1 - step0) * 3 * h->box_floats() : nullptr;
      if (h->form3d()) {
        ceep::Step3DFArgs fa{};
        fa.a = a;
        fa.xcd_order = h->form3d() == 1;
        fa.seg_len = h->march_len;
        for (int k = 0; k < 3; ++k) {
          fa.e_in[k] = h->f[i][k]; fa.h_in[k] = h->f[i][3 + k];
          fa.e_out[k] = h->f[o][k]; fa.h_out[k] = h->f[o][3 + k];
        }
        if (h->form3d() == 3)
          CEEP_HIP(ceep::launch_step3d_march(fa, poles, bb, s));
        else
          CEEP_HIP(ceep::launch_step3d_fused(fa, poles, bb, s));
        continue;
      }
      CEEP_HIP(ceep::launch_h3d(a, bb, s));
      CEEP_HIP(ceep::launch_e3d(a, poles, bb, s));
    }
  }
  return CEEP_OK;
}

// FP64 buffers of the gradient path's forward: fields (ping-pong), ca/cb
// (fdtd_2d_batched.py:152-159, formed in FP64 and kept in FP64; 0 off the E
// interior), the waveform and the traces.
int prepare_f64(ceep_fdtd* h) {
  int rc;
  const size_t nbox = (size_t)h->box_floats();
  for (int i = 0; i < 2; ++i)
    for (int f = 0; f < 3; ++f)
      if (!h->f64[i][f] && (rc = dalloc(&h->f64[i][f], nbox))) return rc;
  if (!h->traces64 && (rc = dalloc(&h->traces64, (size_t)h->B * h->P * h->T))) return rc;
  hipStream_t s = h->stream;
  if (!h->wave64) {
    if ((rc = dalloc(&h->wave64, h->wave_h.size()))) return rc;
    CEEP_HIP(hipMemcpyAsync(h->wave64, h->wave_h.data(), h->wave_h.size() * sizeof(double),
                            hipMemcpyHostToDevice, s));
  }
  if (h->maps64_dirty || !h->ca64) {
    const long long nmap = h->plane;
    std::vector<double> ca(nmap, 0.0), cb(nmap, 0.0);
    const double dt = h->dt;
    for (int a = 1; a < h->M[0] - 1; ++a)
      for (int b = 1; b < h->M[1] - 1; ++b) {
        const long long g = h->full_index(a + h->org, b + h->org, 0);
        const long long o = h->box_index(a, b, 0);
        const double eps = h->eps_r[g] * kEps0;
        const double sv = h->sigma_e[g] * dt / (2.0 * eps);
        ca[o] = (1.0 - sv) / (1.0 + sv);
        cb[o] = (dt / eps) / (1.0 + sv);
      }
    if (!h->ca64 && (rc = dalloc(&h->ca64, (size_t)nmap))) return rc;
    if (!h->cb64 && (rc = dalloc(&h->cb64, (size_t)nmap))) return rc;
    CEEP_HIP(hipMemcpyAsync(h->ca64, ca.data(), nmap * sizeof(double), hipMemcpyHostToDevice, s));
    CEEP_HIP(hipMemcpyAsync(h->cb64, cb.data(), nmap * sizeof(double), hipMemcpyHostToDevice, s));
    CEEP_HIP(hipStreamSynchronize(s));   // the host vectors go out of scope
    h->maps64_dirty = false;
  }
  return CEEP_OK;
}

// Enqueue FP64 steps [step0, step0 + nsteps) from FP64 buffer `start`; with
// e_save, E after step s goes (as FP32) to slot s + 1 - step0.
int launch_range64(ceep_fdtd* h, int start, int step0, int nsteps, float* e_save,
                   int* end_buf) {
  if (h->use_tb()) {   // blocked: 4 steps per launch on 32-row four-wavefront windows
    constexpr int K = 4, R = 32;
    ceep::TB2D64Args a{};
    a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch; a.B = h->B; a.bstride = h->plane;
    a.ca = h->ca64; a.cb = h->cb64;
    a.ch = (h->dt / kMu0) * (1.0 / h->dx);
    a.idx = 1.0 / h->dx;
    a.src_cell = h->src2; a.wave = h->wave64; a.wave_stride_b = h->wave_stride;
    a.T = h->T; a.P = h->P; a.probe_cell = h->probe_cell;
    a.traces = h->traces64; a.traces32 = h->traces;
    a.slot = h->box_floats();
    a.tiles_x = (h->M[0] + (R - 2 * K) - 1) / (R - 2 * K);
    a.tiles_y = (h->M[1] + (64 - 2 * K) - 1) / (64 - 2 * K);
    a.bpw = 2;
    a.ntasks = a.tiles_x * a.tiles_y * ((h->B + a.bpw - 1) / a.bpw);
    a.xcd_order = h->xcd_order;
    int cur = start;
    for (int st = step0; st < step0 + nsteps; st += K) {
      const int o = cur ^ 1;
      a.ez_in = h->f64[cur][0]; a.hx_in = h->f64[cur][1]; a.hy_in = h->f64[cur][2];
      a.ez_out = h->f64[o][0]; a.hx_out = h->f64[o][1]; a.hy_out = h->f64[o][2];
      a.step0 = st;
      a.ksteps = std::min(K, step0 + nsteps - st);
      a.e_save = e_save ? e_save + (size_t)(st + 1 - step0) * h->box_floats() : nullptr;
      CEEP_HIP(ceep::launch_step2d_f64_tb(a, R, K, h->stream));
      cur = o;
    }
    if (end_buf) *end_buf = cur;
    return CEEP_OK;
  }
  if (end_buf) *end_buf = (start + nsteps) & 1;
  ceep::Step2D64Args a{};
  a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch; a.B = h->B; a.bstride = h->plane;
  a.ca = h->ca64; a.cb = h->cb64;
  a.ch = (h->dt / kMu0) * (1.0 / h->dx);
  a.idx = 1.0 / h->dx;
  a.src_cell = h->src2; a.wave = h->wave64; a.wave_stride_b = h->wave_stride;
  a.T = h->T; a.P = h->P;
  a.probe_head = h->probe_head; a.probe_next = h->probe_next;
  a.traces = h->traces64; a.traces32 = h->traces;
  for (int st = step0; st < step0 + nsteps; ++st) {
    const int i = (start + st - step0) & 1, o = i ^ 1;
    a.ez_in = h->f64[i][0]; a.hx_in = h->f64[i][1]; a.hy_in = h->f64[i][2];
    a.ez_out = h->f64[o][0]; a.hx_out = h->f64[o][1]; a.hy_out = h->f64[o][2];
    a.step = st;
    a.e_save = e_save ? e_save + (size_t)(st + 1 - step0) * h->box_floats() : nullptr;
    CEEP_HIP(ceep::launch_step2d_f64(a, h->stream));
  }
  return CEEP_OK;
}

// copy n doubles with the float4 copy kernel (every buffer is 16-byte aligned)
int copy64(const double* src, double* dst, long long n, hipStream_t s) {
  CEEP_HIP(ceep::launch_copy(reinterpret_cast<const float*>(src), reinterpret_cast<float*>(dst),
                             2 * n, s));
  return CEEP_OK;
}

int ready_check(ceep_fdtd* h) {
  if (!h) return fail(CEEP_EINVAL, "null handle");
  if (!h->have_sources) return fail(CEEP_ESTATE, "sources not set (ceep_fdtd_set_sources)");
  if (!h->have_probes) return fail(CEEP_ESTATE, "probes not set (ceep_fdtd_set_probes)");
  return commit_materials(h);
}

// Advance total_steps steps from the current state (the reference's run()
// continues from the fields it finished with; _build starts them at zero).
// Traces are cleared first.  Caller synchronizes.
int run_device(ceep_fdtd* h) {
  h->traces64_valid = false;
  CEEP_HIP(hipMemsetAsync(h->traces, 0, (size_t)h->B * h->P * h->T * sizeof(float), h->stream));
  const int start = h->parity;
  int rc = CEEP_OK;
  if (!h->use_graph) {
    rc = launch_steps(h, start);
  } else {
    if (!h->gexec[start]) {
      // the TX groups' extra streams and events exist before the capture
      // starts: no HIP object is created while the stream captures
      const int nx = std::max(h->tx_streams, h->tx_streams3d) - 1;
      if (nx > 0 && (rc = ensure_streams(h, nx))) return rc;
      hipGraph_t g = nullptr;
      CEEP_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      rc = launch_steps(h, start);
      hipError_t e = hipStreamEndCapture(h->stream, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return fail(CEEP_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&h->gexec[start], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        h->gexec[start] = nullptr;
        return fail(CEEP_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
    }
    CEEP_HIP(hipGraphLaunch(h->gexec[start], h->stream));
  }
  if (rc) return rc;
  h->parity = start ^ (h->launches_per_run() & 1);
  if (h->ndim == 2 && h->has_debye && !h->use_tb() && !h->cpml() && h->parity != start) {
    // the per-step kernel updates J in place in jd[start]; keep J with the fields
    CEEP_HIP(ceep::launch_copy(h->jd[start], h->jd[h->parity], h->box_floats(), h->stream));
  }
  return CEEP_OK;
}

int in_box(const ceep_fdtd* h, const int64_t* p, bool interior) {
  for (int d = 0; d < h->ndim; ++d) {
    const int64_t a = p[d] - h->org;
    if (interior ? (a < 1 || a > h->M[d] - 2) : (a < 0 || a > h->M[d] - 1)) return 0;
  }
  return 1;
}

}  // namespace

extern "C" {

const char* ceep_fdtd_last_error(void) { return g_err.c_str(); }

int ceep_fdtd_abi_version(void) { return 1; }

int ceep_fdtd_create(int ndim, const int64_t* shape, int cpml_n, int box_mode, int batch,
                     int n_probes, int total_steps, double dx, double dt, ceep_fdtd** out) {
  if (!out) return fail(CEEP_EINVAL, "out is null");
  *out = nullptr;
  if (ndim != 2 && ndim != 3) return fail(CEEP_EINVAL, "ndim must be 2 or 3");
  if (!shape) return fail(CEEP_EINVAL, "shape is null");
  if (box_mode != CEEP_BOX_REFERENCE && box_mode != CEEP_BOX_FULL && box_mode != CEEP_BOX_CPML)
    return fail(CEEP_EINVAL, "unknown box mode");
  if (box_mode == CEEP_BOX_FULL && ndim != 2)
    return fail(CEEP_EINVAL, "the full-grid PEC box (run_cpu) exists for 2D only");
  if (box_mode == CEEP_BOX_CPML) {
    if (cpml_n < 1 || cpml_n > ceep::kMaxCpml)
      return fail(CEEP_EINVAL, "CPML thickness must be in [1, 64]");
    // 2D slabs [0, t) and [N-t, N); 3D H slabs [0, t) and [N-1-t, N-1)
    const int64_t need = 2 * (int64_t)cpml_n + (ndim == 3 ? 1 : 0);
    for (int d = 0; d < ndim; ++d)
      if (need > shape[d])
        return fail(CEEP_EINVAL, "the two CPML slabs of an axis must not overlap "
                                 "(2*thickness <= N in 2D, 2*thickness + 1 <= N in 3D)");
  }
  if (batch < 1 || batch > ceep::kMaxBatch) return fail(CEEP_EINVAL, "batch must be in [1, 1024]");
  if (n_probes < 0 || n_probes > 32767) return fail(CEEP_EINVAL, "n_probes must be in [0, 32767]");
  if (total_steps < 1) return fail(CEEP_EINVAL, "total_steps must be >= 1");
  if (!(dx > 0) || !(dt > 0)) return fail(CEEP_EINVAL, "dx and dt must be positive");
  if (box_mode == CEEP_BOX_REFERENCE && cpml_n < 2)
    return fail(CEEP_EINVAL, "cpml_thickness must be >= 2 (the reference's CPML profile "
                             "divides by cpml_thickness - 1)");
  ceep_fdtd* h = new ceep_fdtd();
  h->ndim = ndim; h->box_mode = box_mode; h->n = cpml_n; h->B = batch; h->P = n_probes;
  h->T = total_steps; h->dx = dx; h->dt = dt;
  h->org = box_mode == CEEP_BOX_REFERENCE ? cpml_n : 0;
  h->full_cells = 1;
  for (int d = 0; d < ndim; ++d) {
    h->N[d] = shape[d];
    h->M[d] = box_mode == CEEP_BOX_REFERENCE ? (int)(shape[d] - 2 * cpml_n + 1) : (int)shape[d];
    if (h->M[d] < 3) {
      delete h;
      return fail(CEEP_EINVAL, "grid too small: each axis needs N - 2*cpml_thickness + 1 >= 3");
    }
    h->full_cells *= shape[d];
  }
  const int inner = h->M[ndim - 1];
  h->pitch = (inner + 63) / 64 * 64;
  if (ndim == 2) {
    h->plane = (long long)h->M[0] * h->pitch;
    if ((long long)h->plane > (1LL << 31) - 1) {
      delete h;
      return fail(CEEP_EUNSUPPORTED, "2D box too large for 32-bit cell indices");
    }
  } else {
    h->row_block = (long long)h->M[1] * h->pitch;
    h->plane = h->M[0] * h->row_block;
  }
  h->ch = (float)((dt / kMu0) * (1.0 / dx));
  h->idx = (float)(1.0 / dx);
  h->eps_r.assign(h->full_cells, 1.0);
  h->sigma_e.assign(h->full_cells, 0.0);
  int rc = CEEP_OK;
  {
    void* sp = nullptr;
    const int src = scratch_address(&sp);
    if (src) { delete h; return src; }
    h->scratch = static_cast<unsigned char*>(sp);
  }
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(CEEP_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  const int nf = ndim == 2 ? 3 : 6;
  for (int k = 0; k < nf && !rc; ++k) {   // the two ping-pong sets of a component are adjacent
    rc = dalloc(&h->f[0][k], 2 * h->box_floats());
    if (!rc) h->f[1][k] = h->f[0][k] + h->box_floats();
  }
  for (int k = 0; k < (ndim == 2 ? 4 : 12) && !rc && h->cpml(); ++k) rc = dalloc(&h->psi[k], h->box_floats());
  if (!rc) rc = dalloc(&h->ca, h->plane);
  if (!rc) rc = dalloc(&h->cb, h->plane);
  if (!rc) rc = dalloc(&h->traces, (size_t)batch * (n_probes > 0 ? n_probes : 1) * total_steps);
  if (!rc) rc = dalloc(&h->probe_head, h->plane);
  if (!rc) rc = dalloc(&h->probe_next, n_probes > 0 ? n_probes : 1);
  if (!rc) rc = dalloc(&h->probe_cell, n_probes > 0 ? n_probes : 1);
  if (!rc) rc = dalloc(&h->partial, 1024);
  if (!rc && ndim == 2) rc = dalloc(&h->src2, batch);
  if (!rc && ndim == 3) rc = dalloc(&h->src3, batch);
  if (!rc) rc = zero_state(h);
  if (!rc && hipStreamSynchronize(h->stream) != hipSuccess)
    rc = fail(CEEP_EHIP, "stream synchronize after zeroing failed");
  if (rc) {
    ceep_fdtd_destroy(h);
    return rc;
  }
  *out = h;
  return CEEP_OK;
}

int ceep_fdtd_reset(ceep_fdtd* h) {
  if (!h) return fail(CEEP_EINVAL, "null handle");
  int rc = commit_materials(h);   // (re)allocates the ADE currents for the pole set
  if (rc) return rc;
  if ((rc = zero_state(h))) return rc;
  CEEP_HIP(hipStreamSynchronize(h->stream));
  return CEEP_OK;
}

int ceep_fdtd_destroy(ceep_fdtd* h) {
  if (!h) return CEEP_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->drop_graph();
  h->free_adjoint();
  for (int k = 0; k < 6; ++k) {
    dfree(h->f[0][k]);   // f[1][k] lives in the same allocation
    h->f[1][k] = nullptr;
  }
  for (int k = 0; k < 3; ++k) { dfree(h->jd[k]); dfree(h->jl[0][k]); dfree(h->jl[1][k]); }
  dfree(h->mpack); dfree(h->jpack);
  dfree(h->ca); dfree(h->cb); dfree(h->da); dfree(h->dbi);
  dfree(h->dk1); dfree(h->dk2); dfree(h->lk1); dfree(h->lk2); dfree(h->lk3);
  dfree(h->src2); dfree(h->src3); dfree(h->wave); dfree(h->probe_head); dfree(h->probe_next); dfree(h->probe_cell); dfree(h->probe_cell3);
  dfree(h->traces); dfree(h->partial); dfree(h->tw); dfree(h->S);
  if (h->S_host) { (void)hipHostFree(h->S_host); h->S_host = nullptr; }
  for (int i = 0; i < 2; ++i)
    for (int k = 0; k < 3; ++k) dfree(h->f64[i][k]);
  dfree(h->ca64); dfree(h->cb64); dfree(h->wave64); dfree(h->traces64);
  h->free_reference();
  dfree(h->eps64); dfree(h->sig64); dfree(h->grad64);
  dfree(h->mat_id); dfree(h->mat_table);
  dfree(h->ref_eps64); dfree(h->ref_sig64);
  for (int k = 0; k < 12; ++k) dfree(h->psi[k]);
  for (int i = 0; i < 3; ++i) {
    if (h->xs[i]) (void)hipStreamSynchronize(h->xs[i]);
    if (h->ev_join[i]) (void)hipEventDestroy(h->ev_join[i]);
    if (h->xs[i]) (void)hipStreamDestroy(h->xs[i]);
  }
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CEEP_OK;
}

int ceep_fdtd_set_materials(ceep_fdtd* h, const double* eps_r, const double* sigma_e,
                            const double* mu_r, const double* sigma_m) {
  if (!h || !eps_r || !sigma_e) return fail(CEEP_EINVAL, "null argument");
  for (long long i = 0; i < h->full_cells; ++i)
    if (!(eps_r[i] > 0) || !std::isfinite(sigma_e[i]))
      return fail(CEEP_EINVAL, "eps_r must be > 0 and sigma_e finite everywhere");
  // the same maps again (e.g. a gradient at the medium the handle holds):
  // nothing to rebuild, and the coefficient maps stay as they are
  const size_t nb = (size_t)h->full_cells * sizeof(double);
  auto same = [&](const std::vector<double>& have, const double* in) {
    return have.size() == (size_t)h->full_cells && std::memcmp(have.data(), in, nb) == 0;
  };
  const bool mag_in = h->ndim == 3 && (mu_r || sigma_m);
  if (same(h->eps_r, eps_r) && same(h->sigma_e, sigma_e) && !mag_in && !h->magnetic &&
      h->mu_r.empty())
    return CEEP_OK;
  h->eps_r.assign(eps_r, eps_r + h->full_cells);
  h->sigma_e.assign(sigma_e, sigma_e + h->full_cells);
  h->magnetic = false;
  if (h->ndim == 3 && (mu_r || sigma_m)) {
    h->mu_r.assign(h->full_cells, 1.0);
    h->sigma_m.assign(h->full_cells, 0.0);
    if (mu_r) h->mu_r.assign(mu_r, mu_r + h->full_cells);
    if (sigma_m) h->sigma_m.assign(sigma_m, sigma_m + h->full_cells);
    for (long long i = 0; i < h->full_cells; ++i) {
      if (!(h->mu_r[i] > 0)) return fail(CEEP_EINVAL, "mu_r must be > 0");
      if (h->mu_r[i] != 1.0 || h->sigma_m[i] != 0.0) h->magnetic = true;
    }
  } else if (h->ndim == 2 && (mu_r || sigma_m)) {
    return fail(CEEP_EINVAL, "2D TMz takes no magnetic maps (the reference fixes da=1, db=dt/mu0)");
  }
  h->mats_dirty = true;
  h->maps64_dirty = true;
  return CEEP_OK;
}

int ceep_fdtd_set_debye(ceep_fdtd* h, const double* delta_eps, const double* tau) {
  if (!h || !delta_eps || !tau) return fail(CEEP_EINVAL, "null argument");
  h->deb_de.assign(delta_eps, delta_eps + h->full_cells);
  h->deb_tau.assign(tau, tau + h->full_cells);
  h->has_debye = true;
  h->mats_dirty = true;
  return CEEP_OK;
}

int ceep_fdtd_set_lorentz(ceep_fdtd* h, const double* delta_eps, const double* omega_0,
                          const double* delta) {
  if (!h || !delta_eps || !omega_0 || !delta) return fail(CEEP_EINVAL, "null argument");
  h->lor_de.assign(delta_eps, delta_eps + h->full_cells);
  h->lor_w0.assign(omega_0, omega_0 + h->full_cells);
  h->lor_dl.assign(delta, delta + h->full_cells);
  h->has_lorentz = true;
  h->mats_dirty = true;
  return CEEP_OK;
}

int ceep_fdtd_clear_poles(ceep_fdtd* h) {
  if (!h) return fail(CEEP_EINVAL, "null handle");
  h->has_debye = h->has_lorentz = false;
  h->deb_de.clear(); h->deb_tau.clear(); h->lor_de.clear(); h->lor_w0.clear(); h->lor_dl.clear();
  h->mats_dirty = true;
  return CEEP_OK;
}

int ceep_fdtd_set_sources(ceep_fdtd* h, const int64_t* pos, const double* waveform,
                          int per_batch) {
  if (!h || !pos || !waveform) return fail(CEEP_EINVAL, "null argument");
  for (int b = 0; b < h->B; ++b) {
    const int64_t* p = pos + (size_t)b * h->ndim;
    for (int d = 0; d < h->ndim; ++d)
      if (p[d] < 0 || p[d] >= h->N[d])
        return fail(CEEP_EINVAL, "source " + std::to_string(b) + " lies outside the grid");
    if (!in_box(h, p, h->box_mode == CEEP_BOX_REFERENCE))
      return fail(CEEP_EUNSUPPORTED,
                  "source " + std::to_string(b) +
                      " lies outside the E-update interior [cpml_thickness+1, N-cpml_thickness-1]; "
                      "these kernels step only the cavity the reference's inert CPML encloses");
  }
  h->src.assign(pos, pos + (size_t)h->B * h->ndim);
  const size_t nw = per_batch ? (size_t)h->B * h->T : (size_t)h->T;
  std::vector<float> w(nw);
  for (size_t i = 0; i < nw; ++i) w[i] = (float)waveform[i];
  h->wave_stride = per_batch ? h->T : 0;
  h->wave_h.assign(waveform, waveform + nw);
  dfree(h->wave64);
  dfree(h->wave);
  int rc = dalloc(&h->wave, nw);
  if (rc) return rc;
  if ((rc = upload_f32(h->wave, w, h->stream))) return rc;
  if (h->ndim == 2) {
    std::vector<int> sc(h->B);
    for (int b = 0; b < h->B; ++b)
      sc[b] = (int)h->box_index((int)(h->src[2 * b] - h->org), (int)(h->src[2 * b + 1] - h->org), 0);
    CEEP_HIP(hipMemcpyAsync(h->src2, sc.data(), sc.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  } else {
    std::vector<long long> sc(h->B);
    for (int b = 0; b < h->B; ++b)
      sc[b] = h->box_index((int)(h->src[3 * b] - h->org), (int)(h->src[3 * b + 1] - h->org),
                           (int)(h->src[3 * b + 2] - h->org));
    CEEP_HIP(hipMemcpyAsync(h->src3, sc.data(), sc.size() * sizeof(long long), hipMemcpyHostToDevice, h->stream));
  }
  CEEP_HIP(hipStreamSynchronize(h->stream));
  h->drop_graph();   // the wave buffer may have moved
  h->have_sources = true;
  return CEEP_OK;
}

int ceep_fdtd_set_probes(ceep_fdtd* h, const int64_t* pos) {
  if (!h || (!pos && h->P > 0)) return fail(CEEP_EINVAL, "null argument");
  std::vector<short> head(h->plane, (short)-1), next(h->P > 0 ? h->P : 1, (short)-1);
  for (int q = h->P - 1; q >= 0; --q) {
    const int64_t* p = pos + (size_t)q * h->ndim;
    for (int d = 0; d < h->ndim; ++d)
      if (p[d] < 0 || p[d] >= h->N[d])
        return fail(CEEP_EINVAL, "probe " + std::to_string(q) + " lies outside the grid");
    if (!in_box(h, p, false)) continue;   // stays exactly 0 in the reference
    const long long o = h->ndim == 2
                            ? h->box_index((int)(p[0] - h->org), (int)(p[1] - h->org), 0)
                            : h->box_index((int)(p[0] - h->org), (int)(p[1] - h->org), (int)(p[2] - h->org));
    next[q] = head[o];
    head[o] = (short)q;
  }
  h->prb.assign(pos, pos + (size_t)h->P * h->ndim);
  if (h->ndim == 2 && h->P > 0) {
    std::vector<int> pc(h->P, -1);
    for (int q = 0; q < h->P; ++q) {
      const int64_t* p = pos + (size_t)q * 2;
      if (in_box(h, p, false))
        pc[q] = (int)h->box_index((int)(p[0] - h->org), (int)(p[1] - h->org), 0);
    }
    CEEP_HIP(hipMemcpyAsync(h->probe_cell, pc.data(), pc.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  }
  if (h->ndim == 3 && h->P > 0) {
    std::vector<long long> pc(h->P, -1);
    for (int q = 0; q < h->P; ++q) {
      const int64_t* p = pos + (size_t)q * 3;
      if (in_box(h, p, false))
        pc[q] = h->box_index((int)(p[0] - h->org), (int)(p[1] - h->org), (int)(p[2] - h->org));
    }
    dfree(h->probe_cell3);
    int rc = dalloc(&h->probe_cell3, pc.size());
    if (rc) return rc;
    CEEP_HIP(hipMemcpyAsync(h->probe_cell3, pc.data(), pc.size() * sizeof(long long), hipMemcpyHostToDevice, h->stream));
    CEEP_HIP(hipStreamSynchronize(h->stream));   // pc goes out of scope
  }
  CEEP_HIP(hipMemcpyAsync(h->probe_head, head.data(), head.size() * sizeof(short), hipMemcpyHostToDevice, h->stream));
  CEEP_HIP(hipMemcpyAsync(h->probe_next, next.data(), next.size() * sizeof(short), hipMemcpyHostToDevice, h->stream));
  CEEP_HIP(hipStreamSynchronize(h->stream));
  h->have_probes = true;
  return CEEP_OK;
}

int ceep_fdtd_run(ceep_fdtd* h, double* traces_out) {
  int rc = ready_check(h);
  if (rc) return rc;
  if ((rc = run_device(h))) return rc;
  CEEP_HIP(hipStreamSynchronize(h->stream));
  h->has_run = true;
  if (traces_out) return ceep_fdtd_get_traces(h, traces_out);
  return CEEP_OK;
}

int ceep_fdtd_get_traces(ceep_fdtd* h, double* out) {
  if (!h || !out) return fail(CEEP_EINVAL, "null argument");
  if (!h->has_run) return fail(CEEP_ESTATE, "no run yet");
  const size_t n = (size_t)h->B * h->P * h->T;
  if (h->traces64_valid) {
    if (n) CEEP_HIP(hipMemcpy(out, h->traces64, n * sizeof(double), hipMemcpyDeviceToHost));
    return CEEP_OK;
  }
  std::vector<float> tmp(n);
  if (n) CEEP_HIP(hipMemcpy(tmp.data(), h->traces, n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) out[i] = tmp[i];
  return CEEP_OK;
}

int ceep_fdtd_dft(ceep_fdtd* h, const double* freqs, int n_freq, double* s_out) {
  if (!h || !freqs || !s_out || n_freq < 1) return fail(CEEP_EINVAL, "bad argument");
  if (!h->has_run) return fail(CEEP_ESTATE, "no run yet");
  if (h->P == 0) return CEEP_OK;
  const int F = n_freq, T = h->T;
  int rc;
  // twiddles (cos, -sin)(w_f s dt) in FP64 -> FP32, cached per frequency list
  if (h->tw_freqs.size() != (size_t)F || !std::equal(freqs, freqs + F, h->tw_freqs.begin())) {
    std::vector<float2> tw((size_t)T * F);
    for (int s = 0; s < T; ++s) {
      const double t = s * h->dt;
      for (int f = 0; f < F; ++f) {
        const double w = 2.0 * M_PI * freqs[f];
        tw[(size_t)s * F + f] = make_float2((float)std::cos(w * t), (float)(-std::sin(w * t)));
      }
    }
    dfree(h->tw);
    dfree(h->S);
    if (h->S_host) { (void)hipHostFree(h->S_host); h->S_host = nullptr; }
    h->tw_freqs.clear();
    if ((rc = dalloc(&h->tw, tw.size()))) return rc;
    if ((rc = dalloc(&h->S, (size_t)F * h->P * h->B))) return rc;
    CEEP_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->S_host),
                           (size_t)F * h->P * h->B * sizeof(float2), hipHostMallocDefault));
    CEEP_HIP(hipMemcpyAsync(h->tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice, h->stream));
    CEEP_HIP(hipStreamSynchronize(h->stream));
    h->tw_freqs.assign(freqs, freqs + F);
  }
  const size_t ns = (size_t)F * h->P * h->B;
  CEEP_HIP(ceep::launch_dft(h->traces, h->tw, h->B, h->P, T, F, (float)h->dt, h->S, h->stream));
  CEEP_HIP(hipMemcpyAsync(h->S_host, h->S, ns * sizeof(float2), hipMemcpyDeviceToHost, h->stream));
  CEEP_HIP(hipStreamSynchronize(h->stream));
  const float2* S = h->S_host;
  for (size_t i = 0; i < ns; ++i) { s_out[2 * i] = S[i].x; s_out[2 * i + 1] = S[i].y; }
  return CEEP_OK;
}

int ceep_fdtd_dft_noalloc(ceep_fdtd* h, const double* freqs, int n_freq, double* s_out) {
  if (!h || !freqs || !s_out || n_freq < 1) return fail(CEEP_EINVAL, "bad argument");
  if (!h->scratch) return fail(CEEP_ESTATE, "no DFT scratch on this handle's device");
  const size_t F = (size_t)n_freq, T = (size_t)h->T, ns = F * h->P * h->B;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t off_tw = up(F * sizeof(double));
  const size_t off_s = off_tw + up(T * F * sizeof(float2));
  const size_t off_z = off_s + up(ns * sizeof(float2));
  const size_t need = off_z + ns * 2 * sizeof(double);
  if (need > kScratchBytes)
    return fail(CEEP_EINVAL, "ceep_fdtd_dft_noalloc: " + std::to_string(n_freq) + " frequencies x " +
                                 std::to_string(h->T) + " steps need " + std::to_string(need) +
                                 " B of scratch; the static scratch holds " +
                                 std::to_string(kScratchBytes) + " B (ceep_fdtd_dft has no limit)");
  if (!h->has_run) return fail(CEEP_ESTATE, "no run yet");
  if (h->P == 0) return CEEP_OK;
  hipStream_t st = h->stream;
  double* d_f = reinterpret_cast<double*>(h->scratch);
  float2* tw = reinterpret_cast<float2*>(h->scratch + off_tw);
  float2* S = reinterpret_cast<float2*>(h->scratch + off_s);
  double* Z = reinterpret_cast<double*>(h->scratch + off_z);
  CEEP_HIP(hipMemcpyAsync(d_f, freqs, F * sizeof(double), hipMemcpyHostToDevice, st));
  CEEP_HIP(ceep::launch_twiddles(d_f, n_freq, h->T, h->dt, tw, st));
  CEEP_HIP(ceep::launch_dft(h->traces, tw, h->B, h->P, h->T, n_freq, (float)h->dt, S, st));
  CEEP_HIP(ceep::launch_c2z(S, Z, (long long)ns, st));
  CEEP_HIP(hipMemcpyAsync(s_out, Z, ns * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipStreamSynchronize(st));
  return CEEP_OK;
}

int ceep_fdtd_get_field(ceep_fdtd* h, int comp, double* out) {
  if (!h || !out) return fail(CEEP_EINVAL, "null argument");
  const int nf = h->ndim == 2 ? 3 : 6;
  if (comp < 0 || comp >= nf) return fail(CEEP_EINVAL, "bad field component");
  const size_t total = (size_t)h->B * h->full_cells;
  std::memset(out, 0, total * sizeof(double));
  const int buf = h->fbuf();
  std::vector<float> box(h->box_floats());
  CEEP_HIP(hipMemcpy(box.data(), h->f[buf][comp], box.size() * sizeof(float), hipMemcpyDeviceToHost));
  const int mz = h->ndim == 3 ? h->M[2] : 1;
  for (int b = 0; b < h->B; ++b)
    for (int a = 0; a < h->M[0]; ++a)
      for (int y = 0; y < h->M[1]; ++y)
        for (int z = 0; z < mz; ++z) {
          const long long g = h->full_index(a + h->org, y + h->org, h->ndim == 3 ? z + h->org : 0);
          out[(size_t)b * h->full_cells + g] = box[(size_t)b * h->plane + h->box_index(a, y, z)];
        }
  return CEEP_OK;
}

int ceep_fdtd_energy(ceep_fdtd* h, double* out) {
  if (!h || !out) return fail(CEEP_EINVAL, "null argument");
  *out = 0.0;
  const int buf = h->fbuf();
  const int ne = h->ndim == 2 ? 1 : 3;
  const int nf = h->ndim == 2 ? 3 : 6;
  double se = 0.0, sh = 0.0;
  std::vector<float> part(1024);
  for (int k = 0; k < nf; ++k) {
    CEEP_HIP(ceep::launch_sumsq(h->f[buf][k], h->box_floats(), h->partial, 1024, h->stream));
    CEEP_HIP(hipMemcpyAsync(part.data(), h->partial, 1024 * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    CEEP_HIP(hipStreamSynchronize(h->stream));
    double acc = 0.0;
    for (float v : part) acc += v;
    (k < ne ? se : sh) += acc;
  }
  const double dV = h->ndim == 2 ? h->dx * h->dx : h->dx * h->dx * h->dx;
  *out = (0.5 * kEps0 * se + 0.5 * kMu0 * sh) * dV;
  return CEEP_OK;
}

int ceep_fdtd_time_runs(ceep_fdtd* h, int reps, double* ms_per_run) {
  if (!h || !ms_per_run || reps < 1) return fail(CEEP_EINVAL, "bad argument");
  int rc = ready_check(h);
  if (rc) return rc;
  // events bracket only the time-step launches of each run, not the zeroing
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  for (auto& e : ev) CEEP_HIP(hipEventCreate(&e));
  for (int r = 0; r < reps && !rc; ++r) {
    rc = zero_state(h);
    if (rc) break;
    hipError_t e = hipEventRecord(ev[2 * r], h->stream);
    if (e != hipSuccess) { rc = fail(CEEP_EHIP, "hipEventRecord failed"); break; }
    rc = run_device(h);
    if (!rc && hipEventRecord(ev[2 * r + 1], h->stream) != hipSuccess)
      rc = fail(CEEP_EHIP, "hipEventRecord failed");
  }
  double total = 0.0;
  if (!rc && hipStreamSynchronize(h->stream) != hipSuccess)
    rc = fail(CEEP_EHIP, "stream synchronize failed");
  for (int r = 0; r < reps && !rc; ++r) {
    float ms = 0.f;
    hipError_t e = hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]);
    if (e != hipSuccess) rc = fail(CEEP_EHIP, std::string("hipEventElapsedTime: ") + hipGetErrorString(e));
    total += ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  h->has_run = true;
  *ms_per_run = total / reps;
  return CEEP_OK;
}

}  // extern "C"

namespace {

// The FP64 forward of ceep_fdtd_run_checkpointed, with FP64 traces.  For the
// reversible adjoint it keeps FP32 roundings of the state at each segment end
// (ck[1..nck]); for the recompute form, FP64 checkpoints every ck_every steps
// and the last segment's E (as FP32) in the store.  The final state is also
// left, rounded, in the FP32 fields.
int run_checkpointed64(ceep_fdtd* h, int nck, bool rev, double* traces_out) {
  int rc;
  const long long nbox = h->box_floats();
  const int every = h->ck_every;
  hipStream_t st = h->stream;
  if ((rc = prepare_f64(h))) return rc;
  if (!rev && (int)h->ck64.size() != nck * 3) {
    for (auto& p : h->ck64) dfree(p);
    h->ck64.assign((size_t)nck * 3, nullptr);
    for (auto& p : h->ck64)
      if ((rc = dalloc(&p, (size_t)nbox))) { h->free_adjoint(); return rc; }
  }
  for (int f = 0; f < 3; ++f) CEEP_HIP(ceep::launch_zero(h->f64[0][f], nbox * sizeof(double), st));
  CEEP_HIP(ceep::launch_zero(h->traces64, (size_t)h->B * h->P * h->T * sizeof(double), st));
  int cur = 0;
  for (int k = 0; k < nck; ++k) {
    // reversible adjoint: FP32 copies of the segment-end states (it never
    // reads checkpoint 0); recompute form: FP64 checkpoints of segment starts
    if (rev) {
      if (k > 0)
        for (int f = 0; f < 3; ++f)
          CEEP_HIP(ceep::launch_to_f32(h->f64[cur][f], h->ck[(size_t)k * 3 + f], nbox, st));
    } else {
      for (int f = 0; f < 3; ++f)
        if ((rc = copy64(h->f64[cur][f], h->ck64[(size_t)k * 3 + f], nbox, st))) return rc;
    }
    const int s0 = k * every, n = std::min(every, h->T - s0);
    float* es = nullptr;
    if (k == nck - 1 && !rev) {
      CEEP_HIP(ceep::launch_to_f32(h->f64[cur][0], h->estore, nbox, st));
      es = h->estore;
    }
    if ((rc = launch_range64(h, cur, s0, n, es, &cur))) return rc;
  }
  for (int f = 0; f < 3; ++f) {
    CEEP_HIP(ceep::launch_to_f32(h->f64[cur][f], h->f[0][f], nbox, st));
    if (rev)   // the final state: where the reversible adjoint starts
      CEEP_HIP(ceep::launch_to_f32(h->f64[cur][f], h->ck[(size_t)nck * 3 + f], nbox, st));
  }
  h->parity = 0;
  CEEP_HIP(hipStreamSynchronize(st));
  h->has_run = true;
  h->traces64_valid = true;
  h->ck_valid = true;
  h->ck_rev = rev;
  h->estore_seg = rev ? -1 : nck - 1;
  if (traces_out) return ceep_fdtd_get_traces(h, traces_out);
  return CEEP_OK;
}

// The live 3D state a checkpoint holds: six fields, then the ADE currents
// (Debye J per component, both Lorentz buffers per component).
std::vector<float*> state3d(ceep_fdtd* h, int par) {
  std::vector<float*> v;
  const int fb = h->form3d() ? par : 0;
  for (int k = 0; k < 6; ++k) v.push_back(h->f[fb][k]);
  if (h->tb3d()) {   // set `par` of the packed currents, as nj chunks of box_floats()
    const long long nbox = h->box_floats();
    for (int j = 0; j < h->nj(); ++j) v.push_back(h->jpack + ((long long)par * h->nj() + j) * nbox);
    return v;
  }
  for (int k = 0; k < 3; ++k) {   // J in place (Lorentz roles by step parity)
    if (h->has_debye) v.push_back(h->jd[k]);
    if (h->has_lorentz) { v.push_back(h->jl[0][k]); v.push_back(h->jl[1][k]); }
  }
  return v;
}

// 3D forward with a checkpoint of the whole state every `every` steps and
// the last segment's E (three components) kept in the store.
int run_checkpointed3d(ceep_fdtd* h, int every, double* traces_out) {
  int rc;
  const int nck = (h->T + every - 1) / every;
  const long long nbox = h->box_floats();
  hipStream_t st = h->stream;
  const size_t nstate = state3d(h, 0).size();
  h->ck_valid = false;
  if (h->ck.size() != (size_t)nck * nstate) {
    for (auto& p : h->ck) dfree(p);
    h->ck.assign((size_t)nck * nstate, nullptr);
    for (auto& p : h->ck)
      if ((rc = dalloc(&p, (size_t)nbox))) { h->free_adjoint(); return rc; }
  }
  h->ck_count = nck;
  if (every != h->ck_every) dfree(h->estore);
  h->ck_every = every;
  h->estore_seg = -1;
  h->estore_off = 0;
  h->estore_tail = -1;
  const size_t slot = (size_t)3 * nbox;
  if (!h->estore) {
    int x = 0;
    if (nck >= 2 && h->adjtb3d()) {
      if (h->estore_extra >= 0) {
        x = std::min(h->estore_extra, every - 1);
      } else {
        // what the blocked adjoint still allocates, plus a margin
        const size_t nj = (size_t)h->nj();
        const size_t later = ((h->adjtb ? 0 : 12) + (h->adjtb_j ? 0 : 2 * nj) + (h->g3tb ? 0 : 2)) *
                                 (size_t)nbox * sizeof(float) +
                             (size_t)h->B * h->P * h->T * sizeof(float) + (size_t(4) << 30);
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
          const size_t need = (size_t)(every + 1) * slot * sizeof(float) + later;
          if (fr > need) x = (int)std::min<size_t>((fr - need) / (slot * sizeof(float)), (size_t)every - 1);
        }
      }
    }
    if (dalloc(&h->estore, (size_t)(every + 1 + x) * slot) != CEEP_OK) {
      x = 0;
      if ((rc = dalloc(&h->estore, (size_t)(every + 1) * slot))) {
        h->free_adjoint();
        return rc;
      }
    }
    h->estore_x = x;
  }
  // the tail segment's first m steps run without the store (m a multiple of
  // the launch depth, so the launches fall as in one range)
  const int kd = h->tb3d() ? h->tb3d_k : 1;
  const int m = h->estore_x > 0 ? (every - h->estore_x + kd - 1) / kd * kd : every;
  const bool tail = nck >= 2 && m < every;
  if ((rc = zero_state(h))) return rc;
  CEEP_HIP(hipMemsetAsync(h->traces, 0, (size_t)h->B * h->P * h->T * sizeof(float), st));
  h->traces64_valid = false;
  int cur = 0;
  for (int k = 0; k < nck; ++k) {
    const int s0 = k * every, n = std::min(every, h->T - s0);
    const std::vector<float*> live = state3d(h, cur);
    for (size_t i = 0; i < nstate; ++i)
      CEEP_HIP(ceep::launch_copy(live[i], h->ck[(size_t)k * nstate + i], nbox, st));
    float* es = nullptr;
    if (k == nck - 1) {
      for (int c = 0; c < 3; ++c)
        CEEP_HIP(ceep::launch_copy(live[c], h->estore + (size_t)c * nbox, nbox, st));
      es = h->estore;
    }
    if (tail && k == nck - 2) {
      // E after step s0 + j goes to slot estore_x + j for j > m
      if ((rc = launch_range(h, cur, s0, m, nullptr))) return rc;
      cur = range_end(h, cur, m);
      if ((rc = launch_range(h, cur, s0 + m, n - m, h->estore + (size_t)(h->estore_x + m) * slot)))
        return rc;
      cur = range_end(h, cur, n - m);
      continue;
    }
    if ((rc = launch_range(h, cur, s0, n, es))) return rc;
    cur = range_end(h, cur, n);
  }
  h->parity = cur;
  CEEP_HIP(hipStreamSynchronize(st));
  h->has_run = true;
  h->ck_valid = true;
  h->estore_seg = nck - 1;
  h->estore_tail = tail ? nck - 2 : -1;
  h->estore_m = tail ? m : 0;
  if (traces_out) return ceep_fdtd_get_traces(h, traces_out);
  return CEEP_OK;
}

// sum over the poles of k2 / 2 (dispersive.py:253-299, grid.py:178-184), the
// part of eps_eff that does not depend on eps_r
double eps_add_at(const ceep_fdtd* h, long long g) {
  const double dt = h->dt;
  double add = 0.0;
  if (h->has_debye && h->deb_tau[g] > 0)
    add += ((2.0 * kEps0 * h->deb_de[g] * dt) / (2.0 * h->deb_tau[g] + dt)) / 2.0;
  if (h->has_lorentz && h->lor_w0[g] > 0) {
    const double d = h->lor_dl[g], w02 = h->lor_w0[g] * h->lor_w0[g], dt2 = dt * dt;
    const double denom = 4.0 + 2.0 * d * dt + w02 * dt2;
    add += (2.0 * kEps0 * h->lor_de[g] * w02 * dt2 / denom) / 2.0;
  }
  return add;
}

// dL/d eps_r and dL/d sigma at one cell from the sensitivities to ca and to
// cb (gcb_over_cb = sum of lam * cb * (curl/dx - sum gJ), i.e. dL/dcb * cb),
// through eps = eps_r eps0 + eps_add, s = sigma dt / (2 eps),
// ca = (1 - s)/(1 + s), cb = (dt/eps)/(1 + s)   (fdtd_3d_batched.py:282-290)
void chain_rule(double eps_r, double eps_add, double sigma, double dt, double dL_dca,
                double gcb_over_cb, double* ge, double* gs) {
  const double eps = eps_r * kEps0 + eps_add;
  const double sv = sigma * dt / (2.0 * eps);
  const double cb = (dt / eps) / (1.0 + sv);
  const double dL_dcb = gcb_over_cb / cb;
  const double dca_ds = -2.0 / ((1.0 + sv) * (1.0 + sv));
  const double ds_deps = -sv / eps, ds_dsig = dt / (2.0 * eps);
  const double dcb_deps = -(dt / (eps * eps)) / (1.0 + sv) +
                          (dt / eps) * (-1.0 / ((1.0 + sv) * (1.0 + sv))) * ds_deps;
  const double dcb_dsig = (dt / eps) * (-1.0 / ((1.0 + sv) * (1.0 + sv))) * ds_dsig;
  *ge = kEps0 * (dL_dca * dca_ds * ds_deps + dL_dcb * dcb_deps);
  *gs = dL_dca * dca_ds * ds_dsig + dL_dcb * dcb_dsig;
}

int adjoint3d_abc(ceep_fdtd* h, const std::vector<float>& dtr, std::vector<float>& gca,
                  std::vector<float>& gcb);
int adjoint3d_blocked(ceep_fdtd* h, const std::vector<float>& dtr, std::vector<float>& gca,
                      std::vector<float>& gcb);

int adjoint3d(ceep_fdtd* h, const double* dtraces, double* grad_eps_r, double* grad_sigma) {
  int rc;
  const int B = h->B, P = h->P, T = h->T;
  const size_t ntr = (size_t)B * P * T;
  double mx = 0.0;
  for (size_t i = 0; i < ntr; ++i) mx = std::max(mx, std::fabs(dtraces[i]));
  std::memset(grad_eps_r, 0, (size_t)h->full_cells * sizeof(double));
  std::memset(grad_sigma, 0, (size_t)h->full_cells * sizeof(double));
  if (mx == 0.0 || P == 0) return CEEP_OK;
  // dL/dtrace is linear in the adjoint: a power-of-two scale keeps the FP32
  // adjoint fields far from the subnormal range
  const double scale = std::ldexp(1.0, -std::ilogb(mx));
  std::vector<float> dtr(ntr);
  for (size_t i = 0; i < ntr; ++i) dtr[i] = (float)(dtraces[i] * scale);
  std::vector<float> gca(h->plane), gcb(h->plane);
  if (h->adjtb3d()) {
    if ((rc = adjoint3d_blocked(h, dtr, gca, gcb))) return rc;
  } else if ((rc = adjoint3d_abc(h, dtr, gca, gcb))) {
    return rc;
  }
  hipStream_t st = h->stream;
  if ((rc = zero_state(h))) return rc;   // the recomputation overwrote the fields
  CEEP_HIP(hipStreamSynchronize(st));
  // the chain rule per cell, x slabs split over host threads (each writes its own cells)
  auto slabs = [&](int x0, int x1) {
    for (int x = x0; x < x1; ++x)
      for (int y = 1; y < h->M[1] - 1; ++y)
        for (int z = 1; z < h->M[2] - 1; ++z) {
          const long long o = h->box_index(x, y, z);
          const long long g = h->full_index(x + h->org, y + h->org, z + h->org);
          chain_rule(h->eps_r[g], eps_add_at(h, g), h->sigma_e[g], h->dt, gca[o] / scale,
                     gcb[o] / scale, &grad_eps_r[g], &grad_sigma[g]);
        }
  };
  const int xs = h->M[0] - 2;
  const int nt = std::max(1, std::min({(int)std::thread::hardware_concurrency(), 16, xs}));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t)
    pool.emplace_back(slabs, 1 + (int)((long long)xs * t / nt), 1 + (int)((long long)xs * (t + 1) / nt));
  slabs(1, 1 + xs / nt);
  for (auto& th : pool) th.join();
  return CEEP_OK;
}

// The three-kernel adjoint (A, B, C per reversed step): gradient sums of ca
// and cb into gca / gcb (host, one plane, scaled like dtr).
int adjoint3d_abc(ceep_fdtd* h, const std::vector<float>& dtr, std::vector<float>& gca,
                  std::vector<float>& gcb) {
  int rc;
  const int B = h->B, P = h->P, T = h->T, every = h->ck_every;
  const long long nbox = h->box_floats();
  const size_t ntr = dtr.size();
  const int bb = h->bb >= 8 ? 8 : h->bb >= 4 ? 4 : (h->bb >= 2 ? 2 : 1);
  const int nchunks = (B + bb - 1) / bb;
  const bool use[6] = {true, true, true, h->has_debye, h->has_lorentz, h->has_lorentz};
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 3; ++c)
      if (use[r] && !h->adj3[r][c] && (rc = dalloc(&h->adj3[r][c], (size_t)nbox))) return rc;
  if (!h->g_ca && (rc = dalloc(&h->g_ca, (size_t)nchunks * h->plane))) return rc;
  if (!h->g_cbi && (rc = dalloc(&h->g_cbi, (size_t)nchunks * h->plane))) return rc;
  if (!h->d_dtr && (rc = dalloc(&h->d_dtr, ntr))) return rc;
  hipStream_t st = h->stream;
  CEEP_HIP(hipMemcpyAsync(h->d_dtr, dtr.data(), ntr * sizeof(float), hipMemcpyHostToDevice, st));
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 3; ++c)
      if (use[r] && r != 2) CEEP_HIP(ceep::launch_zero(h->adj3[r][c], (size_t)nbox * sizeof(float), st));
  CEEP_HIP(ceep::launch_zero(h->g_ca, (size_t)nchunks * h->plane * sizeof(float), st));
  CEEP_HIP(ceep::launch_zero(h->g_cbi, (size_t)nchunks * h->plane * sizeof(float), st));
  // lam_z^T = dL/dtrace(T-1) at the probes
  CEEP_HIP(ceep::launch_adj3d_seed(h->adj3[0][2], h->probe_cell3, h->d_dtr, B, P, T, T - 1,
                                   h->plane, st));
  ceep::Adj3DArgs a{};
  a.Mx = h->M[0]; a.My = h->M[1]; a.Mz = h->M[2]; a.pitch = h->pitch; a.B = B;
  a.sx = h->row_block; a.bstride = h->plane;
  for (int c = 0; c < 3; ++c) {
    a.lam[c] = h->adj3[0][c]; a.hbar[c] = h->adj3[1][c]; a.q[c] = h->adj3[2][c];
    a.ad[c] = h->adj3[3][c]; a.al[c] = h->adj3[4][c]; a.al_prev[c] = h->adj3[5][c];
  }
  a.ca = h->ca; a.cb = h->cb; a.da = h->da; a.dbi = h->dbi;
  a.d_k1 = h->dk1; a.d_k2dt = h->dk2; a.l_k1 = h->lk1; a.l_k2dt = h->lk2; a.l_k3 = h->lk3;
  a.ch = h->ch; a.idx = h->idx;
  a.e_cstride = nbox;
  a.src_cell = h->src3; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
  a.T = T; a.P = P; a.dtr = h->d_dtr;
  a.probe_head = h->probe_head; a.probe_next = h->probe_next;
  a.g_ca = h->g_ca; a.g_cbi = h->g_cbi; a.gstride = h->plane;
  const size_t nstate = state3d(h, 0).size();
  const size_t slot = (size_t)3 * nbox;
  for (int k = h->ck_count - 1; k >= 0; --k) {
    const int s0 = k * every, n = std::min(every, T - s0);
    // the one-step forms keep Lorentz J roles by step parity; the blocked
    // form's sets carry their roles, so it restarts from set 0
    const int par = h->tb3d() ? 0 : (s0 & 1);
    const std::vector<float*> live = state3d(h, par);
    // (this form reads segments from slot 0 only)
    if (k != h->estore_seg || h->estore_off != 0) {   // recompute the segment, storing E
      for (size_t i = 0; i < nstate; ++i)
        CEEP_HIP(ceep::launch_copy(h->ck[(size_t)k * nstate + i], live[i], nbox, st));
      for (int c = 0; c < 3; ++c)
        CEEP_HIP(ceep::launch_copy(h->ck[(size_t)k * nstate + c], h->estore + (size_t)c * nbox, nbox, st));
      if ((rc = launch_range(h, par, s0, n, h->estore))) return rc;
      h->estore_seg = k;
      h->estore_off = 0;
    }
    for (int s = s0 + n - 1; s >= s0; --s) {
      a.e_s = h->estore + (size_t)(s - s0) * slot;
      a.e_s1 = h->estore + (size_t)(s + 1 - s0) * slot;
      a.step = s;
      // C of step s + 1 runs inside A of step s (none before the last step)
      CEEP_HIP(ceep::launch_adj3d(a, h->poles(), bb, s < T - 1, st));
      for (int c = 0; c < 3; ++c) std::swap(a.al[c], a.al_prev[c]);   // the new al is in al_prev
    }
  }
  a.step = 0;   // C of the first step, alone
  CEEP_HIP(ceep::launch_adj3d_c(a, bb, st));
  if (nchunks > 1) {   // fold the chunk partials into plane 0 (in place: plane 0 is read first)
    CEEP_HIP(ceep::launch_sum_parts(h->g_ca, nchunks, h->plane, h->plane, h->g_ca, st));
    CEEP_HIP(ceep::launch_sum_parts(h->g_cbi, nchunks, h->plane, h->plane, h->g_cbi, st));
  }
  CEEP_HIP(hipMemcpyAsync(gca.data(), h->g_ca, gca.size() * sizeof(float), hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipMemcpyAsync(gcb.data(), h->g_cbi, gcb.size() * sizeof(float), hipMemcpyDeviceToHost, st));
  return CEEP_OK;
}

// The blocked adjoint: adj3d_k reversed steps per launch (the segment's last
// launch takes what is left), each a full A, B and C, the state in two sets.
int adjoint3d_blocked(ceep_fdtd* h, const std::vector<float>& dtr, std::vector<float>& gca,
                      std::vector<float>& gcb) {
  int rc;
  const int B = h->B, P = h->P, T = h->T, every = h->ck_every;
  const long long nbox = h->box_floats();
  const size_t ntr = dtr.size();
  const int nj = h->nj();
  hipStream_t st = h->stream;
  if (!h->adjtb && (rc = dalloc(&h->adjtb, (size_t)12 * nbox))) return rc;
  if (nj && !h->adjtb_j && (rc = dalloc(&h->adjtb_j, (size_t)2 * nj * nbox))) return rc;
  if (!h->g3tb && (rc = dalloc(&h->g3tb, (size_t)2 * nbox))) return rc;
  if (!h->d_dtr && (rc = dalloc(&h->d_dtr, ntr))) return rc;
  CEEP_HIP(hipMemcpyAsync(h->d_dtr, dtr.data(), ntr * sizeof(float), hipMemcpyHostToDevice, st));
  CEEP_HIP(ceep::launch_zero(h->adjtb, (size_t)12 * nbox * sizeof(float), st));
  if (nj) CEEP_HIP(ceep::launch_zero(h->adjtb_j, (size_t)2 * nj * nbox * sizeof(float), st));
  CEEP_HIP(ceep::launch_zero(h->g3tb, (size_t)2 * nbox * sizeof(float), st));
  ceep::AdjTB3DArgs t{};
  t.Mx = h->M[0]; t.My = h->M[1]; t.Mz = h->M[2]; t.pitch = h->pitch; t.B = B;
  t.sx = h->row_block; t.bstride = h->plane; t.nbox = nbox;
  for (int c = 0; c < 3; ++c) {
    t.lam[c] = h->adjtb + (size_t)2 * c * nbox;
    t.hbar[c] = h->adjtb + (size_t)(6 + 2 * c) * nbox;
  }
  t.jp = nj ? h->adjtb_j : nullptr;
  t.mp = h->mpack;
  t.magnetic = h->da != nullptr;
  if (h->n_mat > 0) {
    t.mat_id = h->mat_id; t.mat_table = reinterpret_cast<const float*>(h->mat_table);
    t.n_mat = h->n_mat;
  }
  t.ch = h->ch; t.idx = h->idx;
  t.src_cell = h->src3; t.wave = h->wave; t.wave_stride_b = h->wave_stride;
  t.T = T; t.P = P; t.dtr = h->d_dtr;
  t.probe_head = h->probe_head; t.probe_next = h->probe_next;
  t.g = h->g3tb; t.gstride = h->plane;
  t.order = h->adj3d_order;
  t.b0 = 0; t.nb = B;
  const std::vector<TxGroup> agroups = tx_groups(B, 1, std::max(1, h->tx_streams3d_adj));
  // lam_z^T = dL/dtrace(T-1) at the probes (set 0)
  CEEP_HIP(ceep::launch_adj3d_seed(t.lam[2], h->probe_cell3, h->d_dtr, B, P, T, T - 1, h->plane, st));
  const size_t nstate = state3d(h, 0).size();
  const size_t slot = (size_t)3 * nbox;
  int set = 0;
  for (int k = h->ck_count - 1; k >= 0; --k) {
    const int s0 = k * every, n = std::min(every, T - s0);
    if (k != h->estore_seg) {   // recompute the segment from its checkpoint, storing E
      // the tail segment: only its first m steps, into the slots after estore_x
      // (the forward pass left the rest behind slot every)
      const bool tail = k == h->estore_tail && h->estore_x > 0;
      const int m = tail ? h->estore_m : n;   // as the checkpointed pass split it
      const int off = tail ? h->estore_x : 0;
      float* base = h->estore + (size_t)off * slot;
      const std::vector<float*> live = state3d(h, 0);
      for (size_t i = 0; i < nstate; ++i)
        CEEP_HIP(ceep::launch_copy(h->ck[(size_t)k * nstate + i], live[i], nbox, st));
      for (int c = 0; c < 3; ++c)
        CEEP_HIP(ceep::launch_copy(h->ck[(size_t)k * nstate + c], base + (size_t)c * nbox, nbox, st));
      if ((rc = launch_range(h, 0, s0, m, base))) return rc;
      h->estore_seg = k;
      h->estore_off = off;
    }
    const float* ebase = h->estore + (size_t)h->estore_off * slot;
    // the segment's reversed steps as TX groups on separate streams (key 29),
    // forked after the recomputation and joined before the next one
    const int nx = (int)agroups.size() - 1;
    if (nx > 0 && (rc = fork_streams(h, st, nx))) return rc;
    int set_end = set;
    for (int gi = 0; gi <= nx; ++gi) {
      t.b0 = agroups[gi].b0;
      t.nb = agroups[gi].nb;
      const hipStream_t gs = gi ? h->xs[gi - 1] : st;
      int gset = set;
      for (int s = s0 + n - 1; s >= s0;) {
        const int K = std::min(h->adj3d_k, s - s0 + 1);
        t.in_set = gset;
        t.step = s;
        t.e_lo = ebase + (size_t)(s - K + 1 - s0) * slot;
        t.tiles_y = (t.My - 1 + (16 - 2 * K) - 1) / (16 - 2 * K);
        t.tiles_z = (t.Mz - 1 + (64 - 2 * K) - 1) / (64 - 2 * K);
        CEEP_HIP(ceep::launch_adj3d_tb(t, K, h->poles(), gs));
        gset ^= 1;
        s -= K;
      }
      set_end = gset;
    }
    set = set_end;
    if (nx > 0 && (rc = join_streams(h, st, nx))) return rc;
  }
  if (B > 1)   // fold the per-TX planes into plane 0 of each (in place: plane 0 is read first)
    for (int c = 0; c < 2; ++c)
      CEEP_HIP(ceep::launch_sum_parts(h->g3tb + (size_t)c * nbox, B, h->plane, h->plane,
                                      h->g3tb + (size_t)c * nbox, st));
  CEEP_HIP(hipMemcpyAsync(gca.data(), h->g3tb, gca.size() * sizeof(float), hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipMemcpyAsync(gcb.data(), h->g3tb + nbox, gcb.size() * sizeof(float),
                          hipMemcpyDeviceToHost, st));
  return CEEP_OK;
}

}  // namespace

namespace {

// The difference form's reference run: the reference medium's forward over
// all T steps with this handle's sources, probes and launch shape, keeping
// E after every step (slot s of e_ref holds E_ref^s; slot 0 = 0), the state
// at every multiple of `every` and the traces.
int adjoint2d_device(ceep_fdtd* h, double scale, double* grad_eps_r, double* grad_sigma);

// The difference form's tile lists: the tiles diff_tile flags, then the
// others (host round trip of ntiles bytes), and the second stream and the
// two events the split launches synchronise with (created once).
int build_tile_lists(ceep_fdtd* h, int ntiles) {
  hipStream_t st = h->stream;
  std::vector<unsigned char> flags((size_t)ntiles);
  CEEP_HIP(hipMemcpyAsync(flags.data(), h->diff_tile, (size_t)ntiles, hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipStreamSynchronize(st));
  std::vector<int> lists;
  lists.reserve((size_t)ntiles);
  for (int t = 0; t < ntiles; ++t)
    if (flags[t]) lists.push_back(t);
  const int nd = (int)lists.size();
  for (int t = 0; t < ntiles; ++t)
    if (!flags[t]) lists.push_back(t);
  if (h->tile_lists && (h->n_diff_tiles + h->n_plain_tiles) != ntiles) {
    dfree(h->tile_lists);
    h->n_diff_tiles = h->n_plain_tiles = 0;
  }
  int rc;
  if (!h->tile_lists && (rc = dalloc(&h->tile_lists, (size_t)ntiles))) return rc;
  CEEP_HIP(hipMemcpyAsync(h->tile_lists, lists.data(), (size_t)ntiles * sizeof(int),
                          hipMemcpyHostToDevice, st));
  CEEP_HIP(hipStreamSynchronize(st));
  h->n_diff_tiles = nd;
  h->n_plain_tiles = ntiles - nd;
  return ensure_streams(h, 1);
}

int run_reference(ceep_fdtd* h, int every) {
  int rc;
  hipStream_t st = h->stream;
  const size_t nbox = (size_t)h->box_floats();
  const int nck = (h->T + every - 1) / every;
  if (!h->e_ref) {
    const size_t n = (size_t)(h->T + 1) * nbox;
    if (dalloc(&h->e_ref, n))
      return fail(CEEP_ENOMEM, "difference form: the reference run's E store (" +
                                   std::to_string(n * sizeof(float)) + " B) does not fit");
  }
  if (h->ref_ck.size() != (size_t)(nck + 1) * 3) {
    for (auto& p : h->ref_ck) dfree(p);
    h->ref_ck.assign((size_t)(nck + 1) * 3, nullptr);
    for (size_t i = 3; i < h->ref_ck.size(); ++i)   // the state at step 0 is zero: not kept
      if ((rc = dalloc(&h->ref_ck[i], nbox))) return rc;
  }
  if ((rc = zero_state(h))) return rc;
  CEEP_HIP(hipMemsetAsync(h->traces, 0, (size_t)h->B * h->P * h->T * sizeof(float), st));
  CEEP_HIP(hipMemsetAsync(h->e_ref, 0, nbox * sizeof(float), st));
  std::swap(h->ca, h->ca_ref);
  std::swap(h->cb, h->cb_ref);
  int cur = 0;
  for (int k = 0; k < nck && !rc; ++k) {
    if (k > 0)
      for (int f = 0; f < 3 && !rc; ++f)
        if (ceep::launch_copy(h->f[cur][f], h->ref_ck[(size_t)k * 3 + f], (long long)nbox, st) != hipSuccess)
          rc = fail(CEEP_EHIP, "reference run: checkpoint copy failed");
    const int s0 = k * every, n = std::min(every, h->T - s0);
    if (!rc) rc = launch_range(h, cur, s0, n, h->e_ref + (size_t)s0 * nbox);
    cur ^= ((n + h->tb_k - 1) / h->tb_k) & 1;
  }
  std::swap(h->ca, h->ca_ref);
  std::swap(h->cb, h->cb_ref);
  if (rc) return rc;
  for (int f = 0; f < 3; ++f)
    CEEP_HIP(ceep::launch_copy(h->f[cur][f], h->ref_ck[(size_t)nck * 3 + f], (long long)nbox, st));
  h->ref_traces.resize((size_t)h->B * h->P * h->T);
  CEEP_HIP(hipMemcpyAsync(h->ref_traces.data(), h->traces, h->ref_traces.size() * sizeof(float),
                          hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipStreamSynchronize(st));
  h->ref_every = every;
  h->ref_dirty = false;
  return CEEP_OK;
}

}  // namespace

extern "C" {

int ceep_fdtd_run_checkpointed(ceep_fdtd* h, int every, double* traces_out) {
  int rc = ready_check(h);
  if (rc) return rc;
  if (h->cpml()) return fail(CEEP_EUNSUPPORTED, "the adjoint covers the reference's cavity modes, not the CPML mode");
  if (every < 1) return fail(CEEP_EINVAL, "checkpoint interval must be >= 1");
  if (h->ndim == 3) return run_checkpointed3d(h, every, traces_out);
  if (h->has_debye || h->has_lorentz)
    return fail(CEEP_EUNSUPPORTED, "the 2D adjoint covers non-dispersive media (3D covers Debye/Lorentz)");
  const int nck = (h->T + every - 1) / every;
  h->ck_valid = false;
  // checkpoint buffers (FP32: one per segment start plus the final state; or
  // FP64 for the FP64 forward) are kept across calls
  const bool rev = h->use_rev();
  // FP32 checkpoints serve the FP32 forward and, rounded from the FP64 state,
  // the reversible adjoint of the FP64 forward
  if ((!h->fwd64 || rev) && h->ck.size() != (size_t)(nck + 1) * 3) {
    for (auto& p : h->ck) dfree(p);
    h->ck.assign((size_t)(nck + 1) * 3, nullptr);
    for (auto& p : h->ck)
      if ((rc = dalloc(&p, h->box_floats()))) { h->free_adjoint(); return rc; }
  }
  h->ck_count = nck;
  if (every != h->ck_every) dfree(h->estore);   // its slot count follows the interval
  h->ck_every = every;
  h->estore_seg = -1;
  // the adjoint starts with the last segment: store its E during this pass so
  // that segment needs no recomputation (the reversible adjoint reads no store)
  if (!rev && !h->estore && (rc = dalloc(&h->estore, (size_t)(every + 1) * h->box_floats())))
    return rc;
  if ((rc = zero_state(h))) return rc;
  CEEP_HIP(hipMemsetAsync(h->traces, 0, (size_t)h->B * h->P * h->T * sizeof(float), h->stream));
  if (h->fwd64) return run_checkpointed64(h, nck, rev, traces_out);
  h->traces64_valid = false;
  int cur = 0;
  for (int k = 0; k < nck; ++k) {
    // the reversible adjoint starts each segment from the checkpoint at its
    // end, so it never reads checkpoint 0 (the zero state)
    if (k > 0 || !rev)
      for (int f = 0; f < 3; ++f)
        CEEP_HIP(ceep::launch_copy(h->f[cur][f], h->ck[(size_t)k * 3 + f], h->box_floats(), h->stream));
    const int s0 = k * every, n = std::min(every, h->T - s0);
    float* es = nullptr;
    if (k == nck - 1 && !rev) {
      CEEP_HIP(ceep::launch_copy(h->f[cur][0], h->estore, h->box_floats(), h->stream));
      es = h->estore;
    }
    if ((rc = launch_range(h, cur, s0, n, es))) return rc;
    const int launches = h->use_tb() ? (n + h->tb_k - 1) / h->tb_k : n;
    cur ^= launches & 1;
  }
  for (int f = 0; f < 3; ++f)   // the final state: where the reversible adjoint starts
    CEEP_HIP(ceep::launch_copy(h->f[cur][f], h->ck[(size_t)nck * 3 + f], h->box_floats(), h->stream));
  h->parity = cur;
  CEEP_HIP(hipStreamSynchronize(h->stream));
  h->has_run = true;
  h->ck_valid = true;
  h->ck_rev = rev;
  h->estore_seg = rev ? -1 : nck - 1;
  if (traces_out) return ceep_fdtd_get_traces(h, traces_out);
  return CEEP_OK;
}

int ceep_fdtd_set_reference(ceep_fdtd* h, const double* eps_r, const double* sigma_e) {
  if (!h) return fail(CEEP_EINVAL, "null handle");
  if (!eps_r && !sigma_e) {
    h->free_reference();
    dfree(h->ref_eps64);
    dfree(h->ref_sig64);
    h->has_ref = false;
    h->mats_dirty = true;
    return CEEP_OK;
  }
  if (!eps_r || !sigma_e) return fail(CEEP_EINVAL, "reference medium: eps_r and sigma_e are both needed");
  if (h->ndim != 2 || h->cpml())
    return fail(CEEP_EUNSUPPORTED, "the difference form covers the 2D cavity modes");
  for (long long i = 0; i < h->full_cells; ++i)
    if (!(eps_r[i] > 0) || !std::isfinite(sigma_e[i]))
      return fail(CEEP_EINVAL, "reference eps_r must be > 0 and sigma_e finite everywhere");
  int rc;
  const size_t nf = (size_t)h->full_cells;
  if (!h->ref_eps64 && (rc = dalloc(&h->ref_eps64, nf))) return rc;
  if (!h->ref_sig64 && (rc = dalloc(&h->ref_sig64, nf))) return rc;
  CEEP_HIP(hipMemcpyAsync(h->ref_eps64, eps_r, nf * sizeof(double), hipMemcpyHostToDevice, h->stream));
  CEEP_HIP(hipMemcpyAsync(h->ref_sig64, sigma_e, nf * sizeof(double), hipMemcpyHostToDevice, h->stream));
  CEEP_HIP(hipStreamSynchronize(h->stream));
  h->has_ref = true;
  h->ref_dirty = true;
  h->mats_dirty = true;
  return CEEP_OK;
}

int ceep_fdtd_run_residual(ceep_fdtd* h, int every, double* residual_out, double* reference_out) {
  int rc = ready_check(h);
  if (rc) return rc;
  if (!h->has_ref) return fail(CEEP_ESTATE, "no reference medium (ceep_fdtd_set_reference)");
  if (h->ndim != 2 || h->cpml() || h->poles())
    return fail(CEEP_EUNSUPPORTED, "the difference form covers 2D non-dispersive cavity runs");
  if (!h->use_rev() || (h->fwd_waves != 4 && h->fwd_waves != 8) || h->fwd64)
    return fail(CEEP_EUNSUPPORTED, "the difference form needs the multi-wavefront blocked FP32 "
                                   "forward and the reversible adjoint");
  if (every < 1) return fail(CEEP_EINVAL, "checkpoint interval must be >= 1");
  hipStream_t st = h->stream;
  const size_t nbox = (size_t)h->box_floats();
  if (h->ref_dirty || h->ref_every != every) {
    h->ck_valid = false;
    if ((rc = run_reference(h, every))) return rc;
  }
  // tiles whose window meets cells where the media differ
  const int R = h->tb_r, K = h->tb_k;
  const int tx = (h->M[0] + (R - 2 * K) - 1) / (R - 2 * K);
  const int ty = (h->M[1] + (64 - 2 * K) - 1) / (64 - 2 * K);
  if (h->diff_tile_n != tx * ty) {
    dfree(h->diff_tile);
    h->diff_tile_n = 0;
    if ((rc = dalloc(&h->diff_tile, (size_t)tx * ty))) return rc;
    h->diff_tile_n = tx * ty;
  }
  CEEP_HIP(ceep::launch_diff_tiles(h->d_a, h->d_b, h->M[0], h->M[1], h->pitch, R, K, tx, ty,
                                   h->diff_tile, st));
  if ((rc = build_tile_lists(h, tx * ty))) return rc;
  const int nck = (h->T + every - 1) / every;
  h->ck_valid = false;
  if (h->ck.size() != (size_t)(nck + 1) * 3) {
    for (auto& p : h->ck) dfree(p);
    h->ck.assign((size_t)(nck + 1) * 3, nullptr);
    for (auto& p : h->ck)
      if ((rc = dalloc(&p, nbox))) { h->free_adjoint(); return rc; }
  }
  h->ck_count = nck;
  if (every != h->ck_every) dfree(h->estore);
  h->ck_every = every;
  h->estore_seg = -1;
  if ((rc = zero_state(h))) return rc;
  CEEP_HIP(hipMemsetAsync(h->traces, 0, (size_t)h->B * h->P * h->T * sizeof(float), st));
  h->traces64_valid = false;
  int cur = 0;
  h->diff_active = true;
  // the adjoint reverses the medium's own total field: each checkpoint is
  // written as E_ref + (E - E_ref), the reference state plus the difference
  for (int k = 0; k < nck && !rc; ++k) {
    if (k > 0)
      for (int f = 0; f < 3 && !rc; ++f)
        if (ceep::launch_add(h->f[cur][f], h->ref_ck[(size_t)k * 3 + f], h->ck[(size_t)k * 3 + f],
                             (long long)nbox, st) != hipSuccess)
          rc = fail(CEEP_EHIP, "difference run: checkpoint failed");
    const int s0 = k * every, n = std::min(every, h->T - s0);
    if (!rc) rc = launch_range(h, cur, s0, n, nullptr);
    cur ^= ((n + h->tb_k - 1) / h->tb_k) & 1;
  }
  h->diff_active = false;
  if (rc) return rc;
  for (int f = 0; f < 3; ++f)
    CEEP_HIP(ceep::launch_add(h->f[cur][f], h->ref_ck[(size_t)nck * 3 + f],
                              h->ck[(size_t)nck * 3 + f], (long long)nbox, st));
  h->parity = cur;
  CEEP_HIP(hipStreamSynchronize(st));
  h->has_run = true;
  h->ck_valid = true;
  h->ck_rev = true;
  if (reference_out)
    for (size_t i = 0; i < h->ref_traces.size(); ++i) reference_out[i] = h->ref_traces[i];
  if (residual_out) return ceep_fdtd_get_traces(h, residual_out);
  return CEEP_OK;
}

int ceep_fdtd_adjoint(ceep_fdtd* h, const double* dtraces, double* grad_eps_r,
                      double* grad_sigma) {
  if (!h || !dtraces || !grad_eps_r || !grad_sigma) return fail(CEEP_EINVAL, "null argument");
  if (!h->ck_valid) return fail(CEEP_ESTATE, "no checkpointed forward run (ceep_fdtd_run_checkpointed)");
  if (h->ndim == 3) return adjoint3d(h, dtraces, grad_eps_r, grad_sigma);
  int rc;
  const int B = h->B, P = h->P, T = h->T;
  // dL/dtrace is linear in the adjoint: scale it by a power of two so the
  // FP32 adjoint fields stay far from the subnormal range, unscale at the end
  double mx = 0.0;
  const size_t ntr = (size_t)B * P * T;
  for (size_t i = 0; i < ntr; ++i) mx = std::max(mx, std::fabs(dtraces[i]));
  std::memset(grad_eps_r, 0, (size_t)h->full_cells * sizeof(double));
  std::memset(grad_sigma, 0, (size_t)h->full_cells * sizeof(double));
  if (mx == 0.0 || P == 0) return CEEP_OK;
  const double scale = std::ldexp(1.0, -std::ilogb(mx));
  if ((rc = pinned(&h->pin_dtr, &h->pin_dtr_n, ntr))) return rc;
  float* dtr = h->pin_dtr;
  for (size_t i = 0; i < ntr; ++i) dtr[i] = (float)(dtraces[i] * scale);
  if (!h->d_dtr && (rc = dalloc(&h->d_dtr, ntr))) return rc;
  CEEP_HIP(hipMemcpyAsync(h->d_dtr, dtr, ntr * sizeof(float), hipMemcpyHostToDevice, h->stream));
  return adjoint2d_device(h, scale, grad_eps_r, grad_sigma);
}

int ceep_fdtd_store_info(ceep_fdtd* h, int64_t* out) {
  if (!h || !out) return fail(CEEP_EINVAL, "null argument");
  out[0] = h->estore ? (int64_t)h->ck_every + 1 + (h->ndim == 3 ? h->estore_x : 0) : 0;
  out[1] = h->ndim == 3 ? h->estore_x : 0;
  out[2] = h->ck_count;
  return CEEP_OK;
}

int ceep_fdtd_get_reference_traces(ceep_fdtd* h, double* out) {
  if (!h || !out) return fail(CEEP_EINVAL, "null argument");
  if (h->ref_traces.empty()) return fail(CEEP_ESTATE, "no reference run yet");
  for (size_t i = 0; i < h->ref_traces.size(); ++i) out[i] = h->ref_traces[i];
  return CEEP_OK;
}

int ceep_fdtd_gradient_misfit(ceep_fdtd* h, int every, const double* freqs, int n_freq,
                              const double* s_target, double* loss_out, double* grad_eps_r,
                              double* grad_sigma) {
  if (!h || !freqs || !loss_out || !grad_eps_r || !grad_sigma || n_freq < 1)
    return fail(CEEP_EINVAL, "bad argument");
  if (h->ndim != 2) return fail(CEEP_EUNSUPPORTED, "the fused misfit gradient covers 2D handles");
  if (h->fwd64) return fail(CEEP_EUNSUPPORTED, "the fused misfit gradient runs the FP32 forward "
                                               "(FP64 forward: ceep_fdtd_adjoint)");
  if (!h->has_ref && !s_target)
    return fail(CEEP_EINVAL, "give S_target or set a reference medium");
  // S_target given: the total-field misfit (a reference medium, if set, is
  // kept for later calls); otherwise the difference form against the reference
  int rc = s_target ? ceep_fdtd_run_checkpointed(h, every, nullptr)
                    : ceep_fdtd_run_residual(h, every, nullptr, nullptr);
  if (rc) return rc;
  const int B = h->B, P = h->P, T = h->T, F = n_freq;
  const size_t nf = (size_t)h->full_cells, nd = (size_t)F * P * B;
  hipStream_t st = h->stream;
  if (P == 0) {
    *loss_out = 0.0;
    std::memset(grad_eps_r, 0, nf * sizeof(double));
    std::memset(grad_sigma, 0, nf * sizeof(double));
    return CEEP_OK;
  }
  // FP64 twiddles, cached per frequency list
  if (h->misfit_freqs.size() != (size_t)F ||
      !std::equal(freqs, freqs + F, h->misfit_freqs.begin())) {
    std::vector<double2> tw((size_t)T * F);
    for (int s = 0; s < T; ++s)
      for (int f = 0; f < F; ++f) {
        const double ph = 2.0 * M_PI * freqs[f] * (s * h->dt);
        tw[(size_t)s * F + f] = make_double2(std::cos(ph), -std::sin(ph));
      }
    dfree(h->tw64);
    dfree(h->misfit_d);
    dfree(h->misfit_tgt);
    h->misfit_freqs.clear();
    if ((rc = dalloc(&h->tw64, tw.size()))) return rc;
    if ((rc = dalloc(&h->misfit_d, nd))) return rc;
    CEEP_HIP(hipMemcpyAsync(h->tw64, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    CEEP_HIP(hipStreamSynchronize(st));
    h->misfit_freqs.assign(freqs, freqs + F);
  }
  if (!h->misfit_out && (rc = dalloc(&h->misfit_out, 2))) return rc;
  const double2* tgt = nullptr;
  if (s_target) {
    if (!h->misfit_tgt && (rc = dalloc(&h->misfit_tgt, nd))) return rc;
    CEEP_HIP(hipMemcpyAsync(h->misfit_tgt, s_target, nd * sizeof(double2), hipMemcpyHostToDevice, st));
    tgt = h->misfit_tgt;
  }
  CEEP_HIP(ceep::launch_misfit_s(h->traces, h->tw64, B, P, T, F, h->dt, tgt, h->misfit_d, st));
  CEEP_HIP(ceep::launch_misfit_sum(h->misfit_d, B, P, F, h->dt, h->misfit_out, st));
  double out[2] = {0.0, 0.0};
  CEEP_HIP(hipMemcpyAsync(out, h->misfit_out, sizeof(out), hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipStreamSynchronize(st));
  *loss_out = out[0];
  if (!(out[1] > 0.0)) {   // the misfit vanishes: so does its gradient
    std::memset(grad_eps_r, 0, nf * sizeof(double));
    std::memset(grad_sigma, 0, nf * sizeof(double));
    return CEEP_OK;
  }
  if (!std::isfinite(out[1])) return fail(CEEP_EINVAL, "the misfit is not finite");
  // dL/dtrace scaled by a power of two at or below 1 / its bound, as
  // ceep_fdtd_adjoint scales the host's dL/dtrace
  const double scale = std::ldexp(1.0, -std::ilogb(out[1]));
  const size_t ntr = (size_t)B * P * T;
  if (!h->d_dtr && (rc = dalloc(&h->d_dtr, ntr))) return rc;
  CEEP_HIP(ceep::launch_misfit_grad(h->misfit_d, h->tw64, B, P, T, F, h->dt, scale, h->d_dtr, st));
  return adjoint2d_device(h, scale, grad_eps_r, grad_sigma);
}

}  // extern "C"

namespace {

// The 2D adjoint from dL/dtrace already in h->d_dtr (FP32, multiplied by
// `scale`, a power of two): reversed segments, gradient sums, chain rule,
// and the two gradient maps copied to the host.
int adjoint2d_device(ceep_fdtd* h, double scale, double* grad_eps_r, double* grad_sigma) {
  int rc;
  const int B = h->B, P = h->P, T = h->T, every = h->ck_every;
  const size_t nbox = (size_t)h->box_floats(), bytes = nbox * sizeof(float);
  const bool rev = h->ck_rev && h->use_rev();
  const bool mw_rev = rev && (h->adj_waves == 4 || h->adj_waves == 8);
  const int adj_k = h->adj_k ? h->adj_k
                             : (mw_rev ? ((h->adj_r == 0 || h->adj_r >= 32) ? 6 : 5) : 4);
  if (h->use_tb() && adj_k > 4 && !mw_rev)
    return fail(CEEP_EINVAL, "adjoint temporal depth above 4 needs the multi-wavefront reversible adjoint");
  if (mw_rev && !ceep::adj2d_rev_mw_shape_ok(h->adj_r ? h->adj_r : 32, adj_k, h->adj_waves,
                                              h->adj_prefetch))
    return fail(CEEP_EINVAL, "reversible adjoint shape: four wavefronts take 16 or 24 rows at depth "
                             "1-5 and 32 at 1-6; eight take 32 or 48 rows at depth 4-6");
  if (!rev && !h->estore && (rc = dalloc(&h->estore, (size_t)(every + 1) * nbox))) return rc;
  for (int i = 0; i < 2; ++i)
    for (int f = 0; f < 3; ++f)
      if (!h->adj[i][f] && (rc = dalloc(&h->adj[i][f], nbox))) return rc;
  // one partial-sum plane per TX chunk of the blocked kernel (at most B)
  const int nchunks = h->use_tb() ? (B + h->adj_bpw - 1) / h->adj_bpw : 1;
  if (!h->g_ca && (rc = dalloc(&h->g_ca, (size_t)B * h->plane))) return rc;
  if (!h->g_cbi && (rc = dalloc(&h->g_cbi, (size_t)B * h->plane))) return rc;
  hipStream_t st = h->stream;
  for (int f = 0; f < 3; ++f) CEEP_HIP(hipMemsetAsync(h->adj[0][f], 0, bytes, st));
  CEEP_HIP(hipMemsetAsync(h->g_ca, 0, (size_t)nchunks * h->plane * sizeof(float), st));
  CEEP_HIP(hipMemsetAsync(h->g_cbi, 0, (size_t)nchunks * h->plane * sizeof(float), st));
  // lam_e^T = dL/dtrace(T-1) at the probes
  CEEP_HIP(ceep::launch_adj_seed(h->adj[0][0], h->probe_cell, h->d_dtr, B, P, T, T - 1,
                                 h->plane, st));
  ceep::Adj2DArgs a{};
  a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch; a.B = B; a.bstride = h->plane;
  a.ca = h->ca; a.cb = h->cb; a.ch = h->ch; a.idx = h->idx;
  a.src_cell = h->src2; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
  a.T = T; a.P = P; a.dtr = h->d_dtr;
  a.probe_head = h->probe_head; a.probe_next = h->probe_next;
  a.g_ca = h->g_ca; a.g_cbi = h->g_cbi;
  int acur = 0;
  // TX groups on separate streams (key 27), as the forward: a group's chunks
  // run every reversed segment on their own stream, fork once, join once
  const std::vector<TxGroup> agroups =
      (rev && h->tx_streams > 1) ? tx_groups(B, h->adj_bpw, h->tx_streams)
                                 : std::vector<TxGroup>{TxGroup{0, B, 0, (B + h->adj_bpw - 1) / h->adj_bpw}};
  const int anx = rev ? (int)agroups.size() - 1 : 0;
  if (anx > 0 && (rc = fork_streams(h, st, anx))) return rc;
  const int acur0 = acur;
  for (size_t gi = 0; gi < agroups.size() && rev; ++gi) {
    const TxGroup& g = agroups[gi];
    const hipStream_t gs = gi ? h->xs[gi - 1] : st;
    const long long po = (long long)g.b0 * h->plane;
    auto at = [po](float* q) { return q + po; };
    acur = acur0;
    for (int k = h->ck_count - 1; k >= 0; --k) {
      // reversible: start from the exact state at the segment's end (the next
      // checkpoint, or the final state) and step forward and adjoint back together
      const int s0 = k * every, n = std::min(every, T - s0);
      const int K = adj_k;
      const int R = h->adj_r ? h->adj_r : (mw_rev ? 32 : 16);
      ceep::AdjTB2DArgs t{};
      t.xcd_order = h->xcd_order;
      t.Mx = h->M[0]; t.My = h->M[1]; t.pitch = h->pitch; t.B = g.nb; t.bstride = h->plane;
      t.ca = h->ca; t.cb = h->cb; t.ch = h->ch; t.idx = h->idx;
      t.src_cell = h->src2 + g.b0; t.wave = h->wave + (long long)g.b0 * h->wave_stride;
      t.wave_stride_b = h->wave_stride;
      t.T = T; t.P = P; t.dtr = h->d_dtr + (size_t)g.b0 * P * T; t.probe_cell = h->probe_cell;
      t.g_ca = h->g_ca + (size_t)g.c0 * h->plane; t.g_cbi = h->g_cbi + (size_t)g.c0 * h->plane;
      t.gstride = h->plane;
      t.tiles_x = (h->M[0] + (R - 2 * K) - 1) / (R - 2 * K);
      t.tiles_y = (h->M[1] + (64 - 2 * K) - 1) / (64 - 2 * K);
      t.bpw = h->adj_bpw;
      t.ntasks = t.tiles_x * t.tiles_y * g.nc;
      // the first launch reads the segment-end checkpoint in place; later ones
      // ping-pong between the two field buffers
      int fcur = 0;
      bool first = true;
      for (int hi = s0 + n; hi > s0; hi -= K) {
        const int o = acur ^ 1, fo = fcur ^ 1;
        t.le_in = at(h->adj[acur][0]); t.mx_in = at(h->adj[acur][1]); t.my_in = at(h->adj[acur][2]);
        t.le_out = at(h->adj[o][0]); t.mx_out = at(h->adj[o][1]); t.my_out = at(h->adj[o][2]);
        float* const* fin = first ? &h->ck[(size_t)(k + 1) * 3] : h->f[fcur];
        t.fe_in = at(fin[0]); t.fhx_in = at(fin[1]); t.fhy_in = at(fin[2]);
        t.fe_out = at(h->f[fo][0]); t.fhx_out = at(h->f[fo][1]); t.fhy_out = at(h->f[fo][2]);
        first = false;
        t.step_hi = hi;
        t.ksteps = std::min(K, hi - s0);
        if (mw_rev)
          CEEP_HIP(ceep::launch_adj2d_rev_mw(t, R, K, h->adj_waves, h->adj_prefetch, gs));
        else
          CEEP_HIP(ceep::launch_adj2d_rev_tb(t, R, K, gs));
        acur = o;
        fcur = fo;
      }
    }
  }
  if (anx > 0 && (rc = join_streams(h, st, anx))) return rc;
  for (int k = h->ck_count - 1; k >= 0 && !rev; --k) {
    const int s0 = k * every, n = std::min(every, T - s0);
    // recompute the segment from its checkpoint, storing E (unless the
    // store already holds this segment)
    if (k != h->estore_seg && h->fwd64) {
      for (int f = 0; f < 3; ++f)
        if ((rc = copy64(h->ck64[(size_t)k * 3 + f], h->f64[0][f], (long long)nbox, st))) return rc;
      CEEP_HIP(ceep::launch_to_f32(h->ck64[(size_t)k * 3], h->estore, (long long)nbox, st));
      if ((rc = launch_range64(h, 0, s0, n, h->estore, nullptr))) return rc;
      h->estore_seg = k;
    } else if (k != h->estore_seg) {
      for (int f = 0; f < 3; ++f)
        CEEP_HIP(ceep::launch_copy(h->ck[(size_t)k * 3 + f], h->f[0][f], (long long)nbox, st));
      CEEP_HIP(ceep::launch_copy(h->ck[(size_t)k * 3], h->estore, (long long)nbox, st));
      if ((rc = launch_range(h, 0, s0, n, h->estore))) return rc;
      h->estore_seg = k;
    }
    if (h->use_tb()) {
      const int K = adj_k;
      const int R = h->adj_r ? h->adj_r : 16;
      ceep::AdjTB2DArgs t{};
      t.xcd_order = h->xcd_order;
      t.Mx = h->M[0]; t.My = h->M[1]; t.pitch = h->pitch; t.B = B; t.bstride = h->plane;
      t.ca = h->ca; t.cb = h->cb; t.ch = h->ch; t.idx = h->idx;
      t.e_store = h->estore; t.slot = (long long)nbox; t.seg0 = s0;
      t.src_cell = h->src2; t.wave = h->wave; t.wave_stride_b = h->wave_stride;
      t.T = T; t.P = P; t.dtr = h->d_dtr; t.probe_cell = h->probe_cell;
      t.g_ca = h->g_ca; t.g_cbi = h->g_cbi; t.gstride = h->plane;
      t.tiles_x = (h->M[0] + (R - 2 * K) - 1) / (R - 2 * K);
      t.tiles_y = (h->M[1] + (64 - 2 * K) - 1) / (64 - 2 * K);
      t.bpw = h->adj_bpw;
      t.ntasks = t.tiles_x * t.tiles_y * ((B + t.bpw - 1) / t.bpw);
      for (int hi = s0 + n; hi > s0; hi -= K) {
        const int o = acur ^ 1;
        t.le_in = h->adj[acur][0]; t.mx_in = h->adj[acur][1]; t.my_in = h->adj[acur][2];
        t.le_out = h->adj[o][0]; t.mx_out = h->adj[o][1]; t.my_out = h->adj[o][2];
        t.step_hi = hi;
        t.ksteps = std::min(K, hi - s0);
        CEEP_HIP(ceep::launch_adj2d_tb(t, R, K, st));
        acur = o;
      }
      continue;
    }
    // adjoint steps s = s0 + n - 1 ... s0
    for (int s = s0 + n - 1; s >= s0; --s) {
      const int o = acur ^ 1;
      a.le_in = h->adj[acur][0]; a.mx_in = h->adj[acur][1]; a.my_in = h->adj[acur][2];
      a.le_out = h->adj[o][0]; a.mx_out = h->adj[o][1]; a.my_out = h->adj[o][2];
      a.e_s = h->estore + (size_t)(s - s0) * nbox;
      a.e_s1 = h->estore + (size_t)(s + 1 - s0) * nbox;
      a.step = s;
      CEEP_HIP(ceep::launch_adj2d(a, st));
      acur = o;
    }
  }
  if (nchunks > 1) {   // fold the chunk partials into plane 0 (in place: plane 0 is read first)
    CEEP_HIP(ceep::launch_sum_parts(h->g_ca, nchunks, h->plane, h->plane, h->g_ca, st));
    CEEP_HIP(ceep::launch_sum_parts(h->g_cbi, nchunks, h->plane, h->plane, h->g_cbi, st));
  }
  // chain rule to eps_r and sigma on the device, in FP64 (ca, cb of
  // fdtd_2d_batched.py:153-157; g_cbi sums lam * (cb / dx) * curl = cb * dL/dcb)
  const size_t nf = (size_t)h->full_cells;
  if (!h->grad64 && (rc = dalloc(&h->grad64, 2 * nf))) return rc;
  ceep::Chain2DArgs c{};
  c.eps_r = h->eps64; c.sigma = h->sig64;
  c.g_ca = h->g_ca; c.g_cbi = h->g_cbi;
  c.inv_scale = 1.0 / scale;   // scale is a power of two: exact
  c.dt = h->dt;
  c.nx = (int)h->N[0]; c.ny = (int)h->N[1]; c.org = h->org;
  c.Mx = h->M[0]; c.My = h->M[1]; c.pitch = h->pitch;
  c.grad_eps_r = h->grad64; c.grad_sigma = h->grad64 + nf;
  CEEP_HIP(ceep::launch_chain2d(c, st));
  // the recomputation overwrote the fields: leave a defined (zero) state
  if ((rc = zero_state(h))) return rc;
  CEEP_HIP(hipMemcpyAsync(grad_eps_r, h->grad64, nf * sizeof(double), hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipMemcpyAsync(grad_sigma, h->grad64 + nf, nf * sizeof(double), hipMemcpyDeviceToHost, st));
  CEEP_HIP(hipStreamSynchronize(st));
  return CEEP_OK;
}

}  // namespace

extern "C" {


int ceep_fdtd_set_tuning(ceep_fdtd* h, int key, int value) {
  if (!h) return fail(CEEP_EINVAL, "null handle");
  if (key == 0) {
    if (value != 1 && value != 2 && value != 4 && value != 8)
      return fail(CEEP_EINVAL, "batch planes per thread must be 1, 2, 4 or 8");
    h->bb = value;
  } else if (key == 1) {
    h->use_graph = value ? 1 : 0;
  } else if (key == 2) {
    if (value < 0 || value > 8) return fail(CEEP_EINVAL, "temporal depth must be in [0, 8]");
    h->tb_k = value;
  } else if (key == 3) {
    if (value != 16 && value != 24 && value != 32 && value != 48 && value != 64)
      return fail(CEEP_EINVAL, "window rows must be 16, 24, 32, 48 or 64");
    h->tb_r = value;
  } else if (key == 4) {
    if (value < 1 || value > 1024) return fail(CEEP_EINVAL, "batch planes per wavefront must be >= 1");
    h->tb_bpw = value;
  } else if (key == 10) {
    if (value != 16 && value != 24 && value != 32 && value != 48)
      return fail(CEEP_EINVAL, "adjoint window rows must be 16, 24, 32 or 48");
    h->adj_r = value;
  } else if (key == 11) {
    if (value < 1 || value > 1024) return fail(CEEP_EINVAL, "adjoint batch planes per wavefront must be >= 1");
    h->adj_bpw = value;
  } else if (key == 14) {
    if (value != 1 && value != 4 && value != 8)
      return fail(CEEP_EINVAL, "adjoint wavefronts per window must be 1, 4 or 8");
    h->adj_waves = value;
  } else if (key == 15) {
    if (value < 1 || value > 6)
      return fail(CEEP_EINVAL, "adjoint temporal depth must be in [1, 6] (5, 6: four-wavefront reversible form only)");
    h->adj_k = value;
  } else if (key == 17) {
    if (value != 1 && value != 2) return fail(CEEP_EINVAL, "3D blocked steps per launch must be 1 or 2");
    h->tb3d_k = value;
  } else if (key == 19) {
    if (value < 0 || value > 2)
      return fail(CEEP_EINVAL, "3D adjoint steps per launch must be 0 (three-kernel form), 1 or 2");
    h->adj3d_k = value;
  } else if (key == 23) {
    if (value < 0 || value > 3) return fail(CEEP_EINVAL, "difference-form tile order must be 0..3");
    h->split_diff = value;
  } else if (key == 24) {
    if (value != 0 && value != 1) return fail(CEEP_EINVAL, "adjoint prefetch must be 0 or 1");
    h->adj_prefetch = value;
  } else if (key == 27) {
    if (value != 1 && value != 2 && value != 4) return fail(CEEP_EINVAL, "TX streams must be 1, 2 or 4");
    h->tx_streams = value;
  } else if (key == 28) {
    if (value != 1 && value != 2 && value != 4) return fail(CEEP_EINVAL, "3D TX streams must be 1, 2 or 4");
    h->tx_streams3d = value;
  } else if (key == 29) {
    if (value != 1 && value != 2 && value != 4)
      return fail(CEEP_EINVAL, "3D adjoint TX streams must be 1, 2 or 4");
    h->tx_streams3d_adj = value;
  } else if (key == 25) {
    if (value < -1) return fail(CEEP_EINVAL, "3D extra E-store slots must be -1 (auto) or >= 0");
    if (value != h->estore_extra) {
      h->estore_extra = value;
      dfree(h->estore);   // re-sized at the next checkpointed run
      h->estore_seg = -1;
      h->estore_off = 0;
      h->estore_x = 0;
      h->estore_tail = -1;
      h->ck_valid = false;
    }
  } else if (key == 22) {
    if (value < 0 || value > 2) return fail(CEEP_EINVAL, "material ids must be 0, 1 or 2");
    if (h->mat_ids_on != value) { h->mat_ids_on = value; h->mats_dirty = true; }
  } else if (key == 20) {
    if (value < 0 || value > 2) return fail(CEEP_EINVAL, "3D adjoint block order must be 0, 1 or 2");
    h->adj3d_order = value;
  } else if (key == 16) {
    if (value != 1 && value != 4 && value != 8)
      return fail(CEEP_EINVAL, "forward wavefronts per window must be 1, 4 or 8");
    h->fwd_waves = value;
  } else if (key == 12) {
    if (value < 0 || value > 3) return fail(CEEP_EINVAL, "3D block order must be 0, 1, 2 or 3");
    h->xcd3d = value;
  } else if (key == 18) {
    if (value < 1 || value > 64) return fail(CEEP_EINVAL, "3D TX group must be in [1, 64]");
    h->tx_group3d = value;
  } else if (key == 9) {
    if (value != 0 && value != 1) return fail(CEEP_EINVAL, "XCD tile order must be 0 or 1");
    h->xcd_order = value;
  } else if (key == 8) {
    if (value < 1 || value > 65536) return fail(CEEP_EINVAL, "x-slabs per marching workgroup must be >= 1");
    h->march_len = value;
  } else if (key == 7) {
    if (value != 0 && value != 1) return fail(CEEP_EINVAL, "adjoint form must be 0 or 1");
    if (value != h->adj_rev) h->ck_valid = false;
    h->adj_rev = value;
  } else if (key == 6) {
    if (value < 0 || value > 4) return fail(CEEP_EINVAL, "3D step form must be 0, 1, 2, 3 or 4");
    // buffer convention: 0 in place, 1-3 ping-pong fields, 4 ping-pong state
    auto conv = [](int f) { return f == 0 ? 0 : f == 4 ? 2 : 1; };
    if (conv(value) != conv(h->fused3d)) {   // the state's buffer convention changes: start over
      h->fused3d = value;
      h->ck_valid = false;
      h->mats_dirty = true;   // the blocked form's packed maps and currents
      int rc = zero_state(h);
      if (rc) return rc;
      CEEP_HIP(hipStreamSynchronize(h->stream));
    }
    h->fused3d = value;
  } else if (key == 5) {
    if (value != 0 && value != 1) return fail(CEEP_EINVAL, "forward precision must be 0 (FP32) or 1 (FP64)");
    if (value && h->ndim != 2) return fail(CEEP_EUNSUPPORTED, "the FP64 forward covers the 2D gradient path");
    if (value != h->fwd64) h->ck_valid = false;
    h->fwd64 = value;
  } else {
    return fail(CEEP_EINVAL, "unknown tuning key");
  }
  h->drop_graph();
  return CEEP_OK;
}

int ceep_fdtd_set_cpml(ceep_fdtd* h, int order, double sigma_factor, double alpha_max,
                       double kappa_max) {
  if (!h) return fail(CEEP_EINVAL, "null handle");
  if (!h->cpml()) return fail(CEEP_ESTATE, "the handle was not created in CEEP_BOX_CPML mode");
  if (order < 0 || order > 16) return fail(CEEP_EINVAL, "CPML grading order must be in [0, 16]");
  if (!(sigma_factor >= 0) || !(alpha_max >= 0) || !(kappa_max >= 1))
    return fail(CEEP_EINVAL, "CPML needs sigma_factor >= 0, alpha_max >= 0, kappa_max >= 1");
  h->cpml_order = order;
  h->cpml_sigma_factor = sigma_factor;
  h->cpml_alpha_max = alpha_max;
  h->cpml_kappa_max = kappa_max;
  h->drop_graph();   // the profiles travel in the launch arguments
  return CEEP_OK;
}

int ceep_beamform(const double* s_matrix, int n_freq, int n_ant, const double* freqs,
                  const int64_t* antennas, const int64_t* x_range, const int64_t* y_range,
                  double dx, double background_permittivity, const double* frequency_range,
                  int method, int normalize, double* image_out) {
  if (!s_matrix || !freqs || !antennas || !x_range || !y_range || !image_out)
    return fail(CEEP_EINVAL, "null argument");
  if (n_freq < 1 || n_ant < 1) return fail(CEEP_EINVAL, "need frequencies and antennas");
  if (method != 0 && method != 1) return fail(CEEP_EINVAL, "method must be 0 (das) or 1 (dmas)");
  if (!(dx > 0) || !(background_permittivity > 0)) return fail(CEEP_EINVAL, "dx and permittivity must be > 0");
  const int64_t nx = x_range[1] - x_range[0], ny = y_range[1] - y_range[0];
  if (nx < 1 || ny < 1 || nx * ny > (1LL << 31)) return fail(CEEP_EINVAL, "bad imaging region");
  // frequency selection and the S columns it keeps (beamforming.py:131-139)
  std::vector<double> fsel;
  std::vector<double> ssel;
  for (int f = 0; f < n_freq; ++f) {
    if (frequency_range && !(freqs[f] >= frequency_range[0] && freqs[f] <= frequency_range[1])) continue;
    fsel.push_back(freqs[f]);
    ssel.insert(ssel.end(), s_matrix + (size_t)f * n_ant * n_ant * 2,
                s_matrix + (size_t)(f + 1) * n_ant * n_ant * 2);
  }
  std::vector<double> ant(2 * (size_t)n_ant);
  for (int k = 0; k < 2 * n_ant; ++k) ant[k] = (double)antennas[k] * dx;
  const size_t npix = (size_t)(nx * ny);
  double *d_s = nullptr, *d_f = nullptr, *d_a = nullptr, *d_o = nullptr;
  int rc = CEEP_OK;
  const size_t nf = fsel.size();
  if (!rc) rc = dalloc(&d_s, std::max<size_t>(ssel.size(), 2));
  if (!rc) rc = dalloc(&d_f, std::max<size_t>(nf, 1));
  if (!rc) rc = dalloc(&d_a, ant.size());
  if (!rc) rc = dalloc(&d_o, npix);
  hipError_t e = hipSuccess;
  if (!rc && nf) e = hipMemcpy(d_s, ssel.data(), ssel.size() * sizeof(double), hipMemcpyHostToDevice);
  if (!rc && nf && e == hipSuccess) e = hipMemcpy(d_f, fsel.data(), nf * sizeof(double), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) e = hipMemcpy(d_a, ant.data(), ant.size() * sizeof(double), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) {
    ceep::BeamformArgs a{};
    a.s = reinterpret_cast<const double2*>(d_s); a.freqs = d_f; a.ant = d_a;
    a.n_freq = (int)nf; a.n_ant = n_ant;
    a.nx = (int)nx; a.ny = (int)ny; a.x0 = (int)x_range[0]; a.y0 = (int)y_range[0];
    a.dx = dx;
    a.velocity = kC0 / std::sqrt(background_permittivity);
    a.two_pi = 2.0 * M_PI;
    a.dmas = method;
    a.out = d_o;
    e = ceep::launch_beamform(a, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(image_out, d_o, npix * sizeof(double), hipMemcpyDeviceToHost);
  }
  dfree(d_s); dfree(d_f); dfree(d_a); dfree(d_o);
  if (rc) return rc;
  if (e != hipSuccess) return fail(CEEP_EHIP, std::string("beamform: ") + hipGetErrorString(e));
  if (normalize) {   // beamforming.py:249-255
    double mn = image_out[0];
    for (size_t k = 1; k < npix; ++k) mn = std::min(mn, image_out[k]);
    double mx = 0.0;
    for (size_t k = 0; k < npix; ++k) { image_out[k] -= mn; mx = std::max(mx, image_out[k]); }
    if (mx > 0)
      for (size_t k = 0; k < npix; ++k) image_out[k] /= mx;
  }
  return CEEP_OK;
}

int ceep_fdtd_box_info(ceep_fdtd* h, int64_t* out) {
  if (!h || !out) return fail(CEEP_EINVAL, "null argument");
  out[0] = h->M[0]; out[1] = h->M[1]; out[2] = h->ndim == 3 ? h->M[2] : 1;
  out[3] = h->pitch; out[4] = h->org;
  return CEEP_OK;
}

}  // extern "C"

// ---- uneven batch: many 2D problems in one persistent launch -----------------

// One TX group of the per-round form: every problem's TX chunks
// [c0, c1) with c0 = chunks * g / G, c1 = chunks * (g + 1) / G (problems with
// no chunk in the group are left out), its own work items enumerated as
// segments (round-major, then problem), launched on its own stream.
struct MultiGroup {
  ceep::MultiDesc* d_desc = nullptr;
  long long* d_seg_start = nullptr;
  int* d_seg_prob = nullptr;
  int* d_seg_round = nullptr;
  int nseg = 0;
  unsigned long long total = 0;
  std::vector<int> round_seg;     // [max_rounds + 1] first segment of each round
  std::vector<long long> seg_start_h;
};

struct ceep_fdtd_multi {
  std::vector<ceep_fdtd*> hs;
  int K = 4, R = 16, bpw = 1, poles = 0, blocks = 0, waves = 1;
  bool mat_ids = false;        // every problem reads material ids (four-wavefront rounds)
  ceep::MultiDesc* d_desc = nullptr;      // whole problems (the persistent launch)
  std::vector<MultiGroup> groups;         // per-round launches, one group per stream
  ceep::PersistItem* d_items = nullptr;   // persistent launch: (problem, TX chunk), largest first
  int nitems = 0;
  unsigned* d_queue = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t xs[3] = {nullptr, nullptr, nullptr};   // the other groups' streams
  hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};   // per mode
  std::vector<int> rounds;
};

namespace {

void multi_free(ceep_fdtd_multi* m) {
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (auto& g : m->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  dfree(m->d_desc);
  for (MultiGroup& g : m->groups) {
    dfree(g.d_desc); dfree(g.d_seg_start); dfree(g.d_seg_prob); dfree(g.d_seg_round);
  }
  dfree(m->d_items); dfree(m->d_queue);
  for (int i = 0; i < 3; ++i) {
    if (m->xs[i]) (void)hipStreamDestroy(m->xs[i]);
    if (m->ev_join[i]) (void)hipEventDestroy(m->ev_join[i]);
  }
  if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

// zero every problem's state and traces and the queue counters, then the
// persistent launch (mode 0) or one launch per round (mode 1)
int multi_enqueue(ceep_fdtd_multi* m, int mode) {
  hipStream_t s = m->stream;
  for (ceep_fdtd* h : m->hs) {
    const size_t bytes = (size_t)h->box_floats() * sizeof(float);
    for (int i = 0; i < 2; ++i)
      for (int f = 0; f < 3; ++f) CEEP_HIP(ceep::launch_zero(h->f[i][f], bytes, s));
    for (int i = 0; i < 2; ++i)
      if (h->jd[i]) CEEP_HIP(ceep::launch_zero(h->jd[i], bytes, s));
    CEEP_HIP(ceep::launch_zero(h->traces, (size_t)h->B * h->P * h->T * sizeof(float), s));
  }
  if (mode == 0) {
    CEEP_HIP(ceep::launch_zero(m->d_queue, sizeof(unsigned), s));
    ceep::PersistArgs p{};
    p.probs = m->d_desc; p.items = m->d_items; p.nitems = m->nitems; p.queue = m->d_queue;
    CEEP_HIP(ceep::launch_persist_mw(p, m->poles, m->R, m->K, m->mat_ids, m->blocks, s));
    return CEEP_OK;
  }
  // the groups share no state: the other streams start after the clears
  // above and rejoin s at the end, so each stream's rounds fill the partly
  // idle last workgroup rounds of the others' launches
  const int ng = (int)m->groups.size();
  if (ng > 1) {
    CEEP_HIP(hipEventRecord(m->ev_fork, s));
    for (int g = 1; g < ng; ++g) CEEP_HIP(hipStreamWaitEvent(m->xs[g - 1], m->ev_fork, 0));
  }
  for (int g = 0; g < ng; ++g) {
    const MultiGroup& G = m->groups[g];
    const hipStream_t gs = g ? m->xs[g - 1] : s;
    ceep::MultiArgs a{};
    a.probs = G.d_desc; a.nprob = (int)m->hs.size();
    a.seg_start = G.d_seg_start; a.seg_prob = G.d_seg_prob; a.seg_round = G.d_seg_round;
    a.nseg = G.nseg; a.total = G.total;
    a.xcd_order = m->hs[0]->xcd_order;   // key 9 of the first problem's handle
    const int nrounds = (int)G.round_seg.size() - 1;
    for (int r = 0; r < nrounds; ++r) {
      const int s0 = G.round_seg[r], s1 = G.round_seg[r + 1];
      if (s1 <= s0) continue;
      const long long first = G.seg_start_h[s0];
      const long long last = s1 < G.nseg ? G.seg_start_h[s1] : (long long)G.total;
      if (m->waves == 4 || m->waves == 8)
        CEEP_HIP(ceep::launch_round_mw(a, m->poles, m->R, m->K, m->waves, s0, s1, first, last - first,
                                       m->mat_ids, gs));
      else
        CEEP_HIP(ceep::launch_round_tb(a, m->poles, m->R, m->K, s0, s1, first, last - first, gs));
    }
  }
  for (int g = 1; g < ng; ++g) {
    CEEP_HIP(hipEventRecord(m->ev_join[g - 1], m->xs[g - 1]));
    CEEP_HIP(hipStreamWaitEvent(s, m->ev_join[g - 1], 0));
  }
  return CEEP_OK;
}

}  // namespace

extern "C" {

int ceep_fdtd_multi_create(int n, ceep_fdtd* const* handles, int K, int R, int bpw, int waves,
                           ceep_fdtd_multi** out) {
  if (!out || !handles || n < 1) return fail(CEEP_EINVAL, "bad argument");
  *out = nullptr;
  if (waves == 1) {
    if (K < 1 || K > 4) return fail(CEEP_EINVAL, "temporal depth must be in [1, 4]");
    if (R != 16 && R != 24 && R != 32) return fail(CEEP_EINVAL, "window rows must be 16, 24 or 32");
  } else if (waves == 4) {
    if (K != 4 && K != 6) return fail(CEEP_EINVAL, "four-wavefront rounds: temporal depth 4 or 6");
    if (R != 32 && R != 48) return fail(CEEP_EINVAL, "four-wavefront rounds: window rows 32 or 48");
  } else if (waves == 8) {
    if ((K != 6 && K != 8) || R != 64)
      return fail(CEEP_EINVAL, "eight-wavefront rounds: depth 6 or 8 on 64-row windows");
  } else {
    return fail(CEEP_EINVAL, "wavefronts per window must be 1, 4 or 8");
  }
  if (bpw < 1) return fail(CEEP_EINVAL, "batch planes per wavefront must be >= 1");
  int rc;
  int poles = -1;
  for (int i = 0; i < n; ++i) {
    ceep_fdtd* h = handles[i];
    if ((rc = ready_check(h))) return rc;
    if (h->ndim != 2 || h->has_lorentz || h->cpml())
      return fail(CEEP_EUNSUPPORTED, "uneven batches take 2D cavity problems without Lorentz poles");
    if (!ceep::tb2d_fits(h->plane))
      return fail(CEEP_EUNSUPPORTED, "uneven batches take 2D boxes of fewer than 2^28 padded cells");
    if (poles >= 0 && poles != h->poles())
      return fail(CEEP_EUNSUPPORTED, "every problem of an uneven batch needs the same pole set");
    poles = h->poles();
  }
  ceep_fdtd_multi* m = new ceep_fdtd_multi();
  m->hs.assign(handles, handles + n);
  m->K = K; m->R = R; m->bpw = bpw; m->poles = poles; m->waves = waves;
  // material ids only when every problem has them (one kernel form per round)
  m->mat_ids = waves == 4 || waves == 8;
  for (int i = 0; i < n; ++i) m->mat_ids = m->mat_ids && handles[i]->n_mat > 0;
  std::vector<ceep::MultiDesc> desc(n);
  int max_rounds = 0;
  for (int i = 0; i < n; ++i) {
    ceep_fdtd* h = handles[i];
    ceep::TB2DArgs& a = desc[i].a;
    a.Mx = h->M[0]; a.My = h->M[1]; a.pitch = h->pitch; a.B = h->B; a.bstride = h->plane;
    a.ca = h->ca; a.cb = h->cb; a.d_k1 = h->dk1; a.d_k2dt = h->dk2;
    if (m->mat_ids) { a.mat_id = h->mat_id; a.mat_table = h->mat_table; a.n_mat = h->n_mat; }
    a.ch = h->ch; a.idx = h->idx;
    a.src_cell = h->src2; a.wave = h->wave; a.wave_stride_b = h->wave_stride;
    a.T = h->T; a.P = h->P; a.probe_cell = h->probe_cell; a.traces = h->traces;
    a.tiles_x = (h->M[0] + (R - 2 * K) - 1) / (R - 2 * K);
    a.tiles_y = (h->M[1] + (64 - 2 * K) - 1) / (64 - 2 * K);
    a.bpw = bpw;
    a.ntasks = a.tiles_x * a.tiles_y * ((h->B + bpw - 1) / bpw);
    for (int b = 0; b < 2; ++b) {
      for (int f = 0; f < 3; ++f) desc[i].f[b][f] = h->f[b][f];
      desc[i].jd[b] = h->jd[b];
    }
    desc[i].nrounds = (h->T + K - 1) / K;
    m->rounds.push_back(desc[i].nrounds);
    max_rounds = std::max(max_rounds, desc[i].nrounds);
  }
  // TX groups of the per-round form (key 27 of the first problem's handle):
  // group g takes chunks [chunks * g / G, chunks * (g + 1) / G) of every
  // problem, its descriptors offset to those planes
  const int G = std::max(1, std::min(4, handles[0]->tx_streams));
  std::vector<std::vector<ceep::MultiDesc>> gdesc;
  std::vector<std::vector<long long>> gseg_start;
  std::vector<std::vector<int>> gseg_prob, gseg_round;
  for (int g = 0; g < G; ++g) {
    std::vector<ceep::MultiDesc> dg = desc;
    std::vector<int> nt(n, 0);
    for (int i = 0; i < n; ++i) {
      ceep_fdtd* h = handles[i];
      const int chunks = (h->B + bpw - 1) / bpw;
      const int c0 = (int)((long long)chunks * g / G), c1 = (int)((long long)chunks * (g + 1) / G);
      if (c1 <= c0) continue;
      const int b0 = c0 * bpw, nb = std::min(h->B, c1 * bpw) - b0;
      const long long po = (long long)b0 * h->plane;
      ceep::TB2DArgs& a = dg[i].a;
      a.B = nb;
      a.ntasks = a.tiles_x * a.tiles_y * (c1 - c0);
      a.src_cell = desc[i].a.src_cell + b0;
      a.wave = desc[i].a.wave + (long long)b0 * desc[i].a.wave_stride_b;
      a.traces = desc[i].a.traces + (long long)b0 * h->P * h->T;
      for (int b = 0; b < 2; ++b) {
        for (int f = 0; f < 3; ++f) dg[i].f[b][f] = desc[i].f[b][f] + po;
        if (desc[i].jd[b]) dg[i].jd[b] = desc[i].jd[b] + po;
      }
      nt[i] = a.ntasks;
    }
    std::vector<long long> seg_start;
    std::vector<int> seg_prob, seg_round;
    unsigned long long cum = 0;
    for (int r = 0; r < max_rounds; ++r)
      for (int i = 0; i < n; ++i)
        if (r < desc[i].nrounds && nt[i] > 0) {
          seg_start.push_back((long long)cum);
          seg_prob.push_back(i);
          seg_round.push_back(r);
          cum += (unsigned long long)nt[i];
        }
    if (cum == 0) continue;              // no problem has a chunk in this group
    MultiGroup mg;
    mg.nseg = (int)seg_start.size();
    mg.total = cum;
    mg.seg_start_h = seg_start;
    for (int k = 0, r = -1; k < mg.nseg; ++k)
      while (r < seg_round[k]) { mg.round_seg.push_back(k); ++r; }
    mg.round_seg.push_back(mg.nseg);
    m->groups.push_back(mg);
    gdesc.push_back(dg);
    gseg_start.push_back(seg_start);
    gseg_prob.push_back(seg_prob);
    gseg_round.push_back(seg_round);
  }
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) { multi_free(m); return fail(CEEP_EHIP, std::string("device query: ") + hipGetErrorString(e)); }
  // the persistent form (four-wavefront windows): one item per (problem, TX
  // chunk), largest first (cells x steps x TX planes), so the last items to
  // start are the shortest; as many 1024-thread workgroups as fit at once
  std::vector<ceep::PersistItem> items;
  std::vector<double> work;
  if (waves == 4) {
    for (int i = 0; i < n; ++i) {
      ceep_fdtd* h = handles[i];
      const int nch = (h->B + bpw - 1) / bpw;
      for (int c = 0; c < nch; ++c) {
        items.push_back(ceep::PersistItem{i, c});
        work.push_back((double)h->M[0] * h->M[1] * h->T * (std::min(h->B, (c + 1) * bpw) - c * bpw));
      }
    }
    std::vector<int> order(items.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    std::vector<ceep::PersistItem> sorted(items.size());
    for (size_t k = 0; k < order.size(); ++k) sorted[k] = items[order[k]];
    items.swap(sorted);
    m->nitems = (int)items.size();
    m->blocks = std::max(1, ceep::persist_mw_blocks_per_cu(poles, R, K, m->mat_ids)) * cus;
  }
  rc = CEEP_OK;
  if (!rc) rc = dalloc(&m->d_desc, (size_t)n);
  for (size_t g = 0; g < m->groups.size() && !rc; ++g) {
    MultiGroup& mg = m->groups[g];
    rc = dalloc(&mg.d_desc, (size_t)n);
    if (!rc) rc = dalloc(&mg.d_seg_start, gseg_start[g].size());
    if (!rc) rc = dalloc(&mg.d_seg_prob, gseg_prob[g].size());
    if (!rc) rc = dalloc(&mg.d_seg_round, gseg_round[g].size());
  }
  if (!rc && m->nitems) rc = dalloc(&m->d_items, items.size());
  if (!rc) rc = dalloc(&m->d_queue, 1);
  if (rc) { multi_free(m); return rc; }
  e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  for (size_t g = 1; g < m->groups.size() && e == hipSuccess; ++g) {
    e = hipStreamCreateWithFlags(&m->xs[g - 1], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_join[g - 1], hipEventDisableTiming);
  }
  if (e == hipSuccess && m->groups.size() > 1) e = hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMemcpy(m->d_desc, desc.data(), desc.size() * sizeof(desc[0]), hipMemcpyHostToDevice);
  for (size_t g = 0; g < m->groups.size() && e == hipSuccess; ++g) {
    MultiGroup& mg = m->groups[g];
    e = hipMemcpy(mg.d_desc, gdesc[g].data(), gdesc[g].size() * sizeof(desc[0]), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(mg.d_seg_start, gseg_start[g].data(), gseg_start[g].size() * sizeof(long long), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(mg.d_seg_prob, gseg_prob[g].data(), gseg_prob[g].size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(mg.d_seg_round, gseg_round[g].data(), gseg_round[g].size() * sizeof(int), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && m->nitems)
    e = hipMemcpy(m->d_items, items.data(), items.size() * sizeof(items[0]), hipMemcpyHostToDevice);
  if (e != hipSuccess) { multi_free(m); return fail(CEEP_EHIP, std::string("multi setup: ") + hipGetErrorString(e)); }
  *out = m;
  return CEEP_OK;
}

int ceep_fdtd_multi_run(ceep_fdtd_multi* m, int mode) {
  if (!m) return fail(CEEP_EINVAL, "null handle");
  if (mode == 2)   // every problem's own run on concurrent streams in one graph
    return fail(CEEP_EUNSUPPORTED, "mode 2 (concurrent runs) is switched off: its first run "
                                   "ended in a host fault whose cause is not found");
  if (mode != 0 && mode != 1) return fail(CEEP_EINVAL, "mode must be 0 (persistent) or 1 (per round)");
  if (mode == 0 && m->waves != 4)
    return fail(CEEP_EUNSUPPORTED, "the persistent launch takes four-wavefront window shapes");
  int rc;
  if (!m->gexec[mode]) {
    hipGraph_t g = nullptr;
    CEEP_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    rc = multi_enqueue(m, mode);
    hipError_t e = hipStreamEndCapture(m->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(CEEP_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&m->gexec[mode], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { m->gexec[mode] = nullptr; return fail(CEEP_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); }
  }
  CEEP_HIP(hipGraphLaunch(m->gexec[mode], m->stream));
  CEEP_HIP(hipStreamSynchronize(m->stream));
  for (size_t i = 0; i < m->hs.size(); ++i) {
    ceep_fdtd* h = m->hs[i];
    h->parity = m->rounds[i] & 1;
    h->has_run = true;
    h->traces64_valid = false;
  }
  return CEEP_OK;
}

int ceep_fdtd_multi_time(ceep_fdtd_multi* m, int mode, int reps, double* ms) {
  if (!m || !ms || reps < 1) return fail(CEEP_EINVAL, "bad argument");
  int rc = ceep_fdtd_multi_run(m, mode);       // capture + warm
  if (rc) return rc;
  hipEvent_t e0, e1;
  CEEP_HIP(hipEventCreate(&e0));
  CEEP_HIP(hipEventCreate(&e1));
  CEEP_HIP(hipEventRecord(e0, m->stream));
  for (int r = 0; r < reps; ++r) CEEP_HIP(hipGraphLaunch(m->gexec[mode], m->stream));
  CEEP_HIP(hipEventRecord(e1, m->stream));
  CEEP_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  CEEP_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / reps;
  return CEEP_OK;
}

int ceep_fdtd_multi_info(ceep_fdtd_multi* m, long long* out) {
  if (!m || !out) return fail(CEEP_EINVAL, "null argument");
  unsigned long long total = 0;
  int nseg = 0;
  for (const MultiGroup& g : m->groups) { total += g.total; nseg += g.nseg; }
  out[0] = (long long)total; out[1] = nseg; out[2] = m->blocks; out[3] = m->nitems;
  return CEEP_OK;
}

int ceep_fdtd_multi_destroy(ceep_fdtd_multi* m) {
  if (m) multi_free(m);
  return CEEP_OK;
}

}  // extern "C"
