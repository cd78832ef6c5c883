// Launch interface of the DAS / DMAS imaging kernel (device pointers, FP64).
#pragma once
#include <hip/hip_runtime.h>

namespace ceep {

struct BeamformArgs {
  const double2* s;       // selected frequencies only: [F][RX][TX] complex
  const double* freqs;    // [F] Hz
  const double* ant;      // [N][2] antenna positions in metres
  int n_freq, n_ant;
  int nx, ny, x0, y0;     // pixel (i, j) sits at ((x0 + i) dx, (y0 + j) dx)
  double dx, velocity, two_pi;
  int dmas;
  double* out;            // [nx][ny] |image|
};

hipError_t launch_beamform(const BeamformArgs& a, hipStream_t s);

}  // namespace ceep
