// Readers of a level operator's planes for the V-cycle kernels (smoother gathers, residual):
// FP32 planes, or FP16 pairs (precision BD_PRECISION_FP16, coarse levels: planes 2w and 2w + 1 of the FP32 layout
// in the low and high half of word w, each value times the level's scale 2^e, e chosen so that the
// largest free diagonal entry lands in (2^13, 2^14]). Blocks start on even planes (6k), so a block
// is three words; the fine along-line scalars EE (30) and NN (31) share word 15.
#pragma once

#include <cstdint>

namespace bd {

struct OpF32 {
  const float* op;
  int Np;
  __device__ __forceinline__ void blk(int base, int src, float o[6]) const {
#pragma unroll
    for (int q = 0; q < 6; q++) o[q] = op[(base + q) * Np + src];
  }
  __device__ __forceinline__ float scalar(int plane, int src) const { return op[plane * Np + src]; }
  __device__ __forceinline__ OpF32 live() const { return *this; }
};

__device__ __forceinline__ float half_lo(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xFFFFu)); }
__device__ __forceinline__ float half_hi(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16)); }

struct OpF16 {
  const uint32_t* op;
  int Np;
  float inv;  // 2^-e
  __device__ __forceinline__ void blk(int base, int src, float o[6]) const {
#pragma unroll
    for (int w = 0; w < 3; w++) {
      const uint32_t v = op[((base >> 1) + w) * Np + src];
      o[2 * w] = half_lo(v) * inv;
      o[2 * w + 1] = half_hi(v) * inv;
    }
  }
  __device__ __forceinline__ float scalar(int plane, int src) const {
    const uint32_t v = op[(plane >> 1) * Np + src];
    return ((plane & 1) ? half_hi(v) : half_lo(v)) * inv;
  }
};

// kernel argument form of OpF16: the inverse scale is computed on the device (hs[1]) and read once
// per thread by live()
struct OpF16Dev {
  const uint32_t* op;
  int Np;
  const float* hs;
  __device__ __forceinline__ OpF16 live() const { return OpF16{op, Np, hs[1]}; }
};

}  // namespace bd
