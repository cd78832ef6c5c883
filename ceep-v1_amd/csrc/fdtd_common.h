// Shared device-side definitions of the batched FDTD engine (gfx950 / CDNA4).
//
// Data layout in HBM (all FP32):
//   field[b][a][c]      2D: b = TX (batch), a = x - n, c = y - n        (c fastest)
//   field[b][a][c][d]   3D: ... d = z - n                               (d fastest)
// over the "cavity box" [n, N-n] on every axis.  With every source strictly
// inside the E-update interior, the reference's CPML auxiliaries and every
// field outside that box stay exactly zero (the profile grading makes
// c[n-1] = b[n-1] - 1 = 0), so only the box is stored and stepped.
// Rows are padded to `pitch` floats (a multiple of 64 -> 256 B aligned rows).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ceep {

constexpr int kMaxPoles = 2;          // Debye + Lorentz per cell
constexpr int kMaxBatch = 1024;

// One ADE pole's per-cell coefficient maps (FP32, cavity box layout without
// the batch axis).  k2 is pre-divided by dt: J += k2dt * (E_new - E_old).
struct PoleMaps {
  const float* k1;
  const float* k2dt;
  const float* k3;     // Lorentz only (J_prev coefficient); null for Debye
};

}  // namespace ceep
