// Launch interface of the 3D step kernels (device pointers, FP32).
#pragma once
#include <hip/hip_runtime.h>

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
};

hipError_t launch_h3d(const Step3DArgs& a, int bb, hipStream_t s);
hipError_t launch_e3d(const Step3DArgs& a, int poles, int bb, hipStream_t s);

}  // namespace ceep
