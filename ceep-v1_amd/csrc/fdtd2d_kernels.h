// Launch interface of the 2D step kernels (device pointers, FP32).
#pragma once
#include <hip/hip_runtime.h>

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
};

// poles: bit 0 = Debye pole present, bit 1 = Lorentz pole present.
// bb: batch planes per thread (1, 2, 4, 8).
hipError_t launch_step2d(const Step2DArgs& a, int poles, int bb, hipStream_t s);

}  // namespace ceep
