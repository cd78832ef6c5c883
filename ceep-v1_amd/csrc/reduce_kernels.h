#pragma once
#include <hip/hip_runtime.h>

namespace ceep {

// traces [B][P][T]; tw [T][F] = (cos(w_f s dt), -sin(w_f s dt)); S [F][P][B].
hipError_t launch_dft(const float* traces, const float2* tw, int B, int P, int T, int F,
                      float dt, float2* S, hipStream_t s);

// dst[0:n) = src[0:n) on stream s (float4 kernel when n % 4 == 0).
hipError_t launch_copy(const float* src, float* dst, long long n, hipStream_t s);

// Zero `bytes` (a multiple of 4) with a kernel, so the clear can sit in a
// captured graph without a memset node.
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);

// dst[0:n) = (float)src[0:n) on stream s.
hipError_t launch_to_f32(const double* src, float* dst, long long n, hipStream_t s);

// partial[nblocks] = per-workgroup sums of x[i]^2 over n floats.
hipError_t launch_sumsq(const float* x, long long n, float* partial, int nblocks,
                        hipStream_t s);

// tw[s][f] = (cos(2 pi freqs[f] s dt), -sin(...)) for s < T, f < F (freqs on the device).
hipError_t launch_twiddles(const double* freqs, int F, int T, double dt, float2* tw,
                           hipStream_t s);

// out = interleaved (re, im) float64 of S[0:n)
hipError_t launch_c2z(const float2* S, double* out, long long n, hipStream_t s);

}  // namespace ceep
