// Device-side coefficient maps, difference-form maps and gradient chain rule
// of the 2D non-dispersive path (coef2d.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace ceep {

struct Coef2DArgs {
  const double* eps_r;     // full grid [nx][ny], FP64
  const double* sigma;
  const double* eps_ref;   // optional reference medium (difference form), or null
  const double* sigma_ref;
  long long ny_full;
  int org;                 // grid index of box cell 0
  int Mx, My, pitch;
  int cpml_interior;       // 1: E interior [1:, 1:] (physical-CPML mode); 0: [1, M-2]
  double dt;
  float* ca;               // box maps [Mx][pitch]
  float* cb;
  float* ca_ref;           // with eps_ref: the reference medium's maps ...
  float* cb_ref;
  float* d_a;              // ... and d_a = ca - cb ca_ref / cb_ref, d_b = cb / cb_ref - 1
  float* d_b;
};
hipError_t launch_coef2d(const Coef2DArgs& a, hipStream_t s);

// flag[t] for every tile of a four-wavefront blocked launch (R rows, depth K):
// 1 if d_a or d_b is nonzero in its window.
hipError_t launch_diff_tiles(const float* d_a, const float* d_b, int Mx, int My, int pitch,
                             int R, int K, int tiles_x, int tiles_y, unsigned char* flag,
                             hipStream_t s);

struct Chain2DArgs {
  const double* eps_r;     // full grid, FP64
  const double* sigma;
  const float* g_ca;       // box plane: dL/dca / inv_scale
  const float* g_cbi;      // box plane: cb dL/dcb / inv_scale
  double inv_scale, dt;
  int nx, ny, org, Mx, My, pitch;
  double* grad_eps_r;      // full grid outputs
  double* grad_sigma;
};
hipError_t launch_chain2d(const Chain2DArgs& a, hipStream_t s);

// out[0:n) = x[0:n) + y[0:n)
hipError_t launch_add(const float* x, const float* y, float* out, long long n, hipStream_t s);

}  // namespace ceep
