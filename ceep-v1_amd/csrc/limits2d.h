// Size limit of the 2D blocked kernels (no HIP dependency, so host code and
// host-side tests can use it).
#pragma once

namespace ceep {

// The 2D blocked kernels (FP32 four- and one-wavefront forms, the reversible
// and store adjoints, the uneven rounds, the FP64 forward) address one TX
// plane with 32-bit byte offsets and mark lanes outside the window with the
// offset 2^31, so every byte offset into a plane -- 8 bytes per cell in the
// FP64 form -- must stay below 2^31: planes (rows x padded columns) of fewer
// than 2^28 cells.  Larger planes take the one-step kernels, which index
// with 64 bits.
constexpr long long kTb2dMaxPlane = 1LL << 28;
constexpr bool tb2d_fits(long long plane_cells) {
  return plane_cells > 0 && plane_cells < kTb2dMaxPlane;
}

}  // namespace ceep
