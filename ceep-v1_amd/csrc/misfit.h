#pragma once
#include <hip/hip_runtime.h>

namespace ceep {

// S-parameter misfit of the gradient path, on the device (FP64 sums).
// traces [B][P][T] (FP32); tw64 [T][F] = (cos(w_f s dt), -sin(w_f s dt)) in FP64;
// s_tgt [F][P][B] or null; D [F][P][B].
//   D = dt sum_s traces[b][p][s] tw64[s][f] - s_tgt
hipError_t launch_misfit_s(const float* traces, const double2* tw64, int B, int P, int T, int F,
                           double dt, const double2* s_tgt, double2* D, hipStream_t s);
// out[0] = sum |D|^2 (the loss); out[1] = max over (p, b) of
// 2 dt sum_f (|Re D| + |Im D|), a bound on |dL/dtrace| (one workgroup).
hipError_t launch_misfit_sum(const double2* D, int B, int P, int F, double dt, double* out,
                             hipStream_t s);
// dtr[b][p][s] = (float)(scale * 2 dt sum_f (Re D tw.x + Im D tw.y)) = scale * dL/dtrace.
hipError_t launch_misfit_grad(const double2* D, const double2* tw64, int B, int P, int T, int F,
                              double dt, double scale, float* dtr, hipStream_t s);

}  // namespace ceep
