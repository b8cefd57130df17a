// Fused forward of the learned-safety network for one wavefront: 64 points,
// one per lane, evaluated with v_mfma_f32_32x32x16_bf16.  Every lane of the
// wave must call it together.
//
// Products are computed transposed, Z (units x points) = W (units x inputs) *
// H (inputs x points), 32 points at a time (two halves of the wave).  The
// matrix-instruction layouts (gfx950):
//   A (32x16):  lane l holds A[l%32][8(l/32) + j], j = 0..7
//   B (16x32):  lane l holds B[8(l/32) + j][l%32]
//   D (32x32):  lane l, register v holds D[8(v/4) + 4(l/32) + v%4][l%32]
// A lane therefore holds units 8g + 4h + i (g, i = 0..3, h = l/32) of its
// point's output, while the next product wants units 8h + j (j = 0..7) of a
// 16-unit block.  Between layers the activations go through a per-wave LDS
// buffer of 32 points x 128 units (BF16): each lane writes its 4-unit runs
// and reads its 8-unit runs back.  Weights (MlpImage) are read through the
// caches; a network is shared by every wave that serves its solid.
#pragma once
#include "mlp.h"
#include "vgm_math.hpp"

namespace vgm {

typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x16 __attribute__((ext_vector_type(16)));

// Normalised coordinate of a point along one axis of the domain box.
VGM_DEV float mlp_coord(double p, double c, double inv_hw) { return float((p - c) * inv_hw); }

// Feature i (0..63) of the point with normalised coordinates u; 0 beyond 50.
// sin(2^k pi u) = sin(2 pi * 2^(k-1) u): the argument in revolutions; the
// cosine is the sine a quarter revolution on.
__device__ __forceinline__ float mlp_feature(int i, float ux, float uy, float uz)
{
  if (i < 3) return i == 0 ? ux : (i == 1 ? uy : uz);
  if (i >= kMlpIn) return 0.0f;
  const int k = (i - 3) / 6, r = (i - 3) % 6, a = r % 3;
  const float u = a == 0 ? ux : (a == 1 ? uy : uz);
  const float rev = u * float(1 << k) * 0.5f;
  return __builtin_amdgcn_sinf(r < 3 ? rev : rev + 0.25f);
}

// SiLU with a hardware reciprocal (one rounding more than a division).
__device__ __forceinline__ float mlp_silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

// Bias of the 16 accumulator rows of lane half h in unit block mb, as the
// accumulator's initial value: rows 32mb + 8g + 4h + i, i = 0..3, g = 0..3.
__device__ __forceinline__ mlp_f32x16 mlp_bias_acc(const float *b, int mb, int h)
{
  mlp_f32x16 acc;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const mlp_f32x4 v = *reinterpret_cast<const mlp_f32x4 *>(b + 32 * mb + 8 * g + 4 * h);
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[4 * g + i] = v[i];
  }
  return acc;
}

// Per-wave activation buffer: 32 points x 128 units, rows padded by 16 bytes.
typedef __bf16 mlp_act_rows[32][kMlpRowH];

// LDS traffic between lanes of one wave: wait for this lane's LDS operations
// and keep the compiler from moving LDS accesses across the point.
__device__ __forceinline__ void mlp_wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// acc = bias + W[32mb.., :] * input, input blocks from `in` (registers, NK
// blocks) or from the activation buffer (NK = 8).  One accumulation chain:
// the waves of the workgroup overlap each other's matrix-instruction latency.
template <int NK, int ROW, bool FromLds>
__device__ __forceinline__ mlp_f32x16 mlp_block(const __bf16 (*w)[ROW], const float *b, const mlp_bf16x8 *in,
                                                const mlp_act_rows &act, int mb, int lane)
{
  const int r32 = lane & 31, h = lane >> 5;
  mlp_f32x16 acc = mlp_bias_acc(b, mb, h);
#pragma unroll
  for (int kb = 0; kb < NK; kb += 2) {
    const mlp_bf16x8 a0 = *reinterpret_cast<const mlp_bf16x8 *>(&w[32 * mb + r32][16 * kb + 8 * h]);
    const mlp_bf16x8 a1 = *reinterpret_cast<const mlp_bf16x8 *>(&w[32 * mb + r32][16 * (kb + 1) + 8 * h]);
    mlp_bf16x8 b0, b1;
    if constexpr (FromLds) {
      b0 = *reinterpret_cast<const mlp_bf16x8 *>(&act[r32][16 * kb + 8 * h]);
      b1 = *reinterpret_cast<const mlp_bf16x8 *>(&act[r32][16 * (kb + 1) + 8 * h]);
    } else {
      b0 = in[kb];
      b1 = in[kb + 1];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);  // operands of one pair of blocks in flight at a time
  }
  return acc;
}

typedef __bf16 mlp_bf16x4 __attribute__((ext_vector_type(4)));

// SiLU of a block's accumulators, BF16, as 4 runs of 4 units (8g + 4h + i).
__device__ __forceinline__ void mlp_act_block(const mlp_f32x16 &acc, mlp_bf16x4 *o)
{
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int i = 0; i < 4; ++i) o[g][i] = static_cast<__bf16>(mlp_silu(acc[4 * g + i]));
}

// Network output y for the point of each lane (normalised coordinates u);
// `act` is this wave's activation buffer.
__device__ __forceinline__ float mlp_eval_wave(const MlpImage &m, mlp_act_rows &act, float ux, float uy, float uz)
{
  const int lane = int(threadIdx.x & 63u);
  const int r32 = lane & 31, h = lane >> 5;
  float mine = 0.0f;
#pragma unroll 1
  for (int nb = 0; nb < 2; ++nb) {
    const int src = 32 * nb + r32;
    const float px = __shfl(ux, src), py = __shfl(uy, src), pz = __shfl(uz, src);
    mlp_bf16x8 x0[4];  // this lane's features 16kb + 8h + j
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int j = 0; j < 8; ++j) x0[kb][j] = static_cast<__bf16>(mlp_feature(16 * kb + 8 * h + j, px, py, pz));
    mlp_bf16x4 o[4][4];  // [mb][g]: this lane's outputs of the current layer
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      mlp_act_block(mlp_block<4, kMlpRow1, false>(m.w1, m.b1, x0, act, mb, lane), o[mb]);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll 1
    for (int l = 0; l < 3; ++l) {
      mlp_wave_sync();  // previous reads of the buffer are done
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int g = 0; g < 4; ++g) *reinterpret_cast<mlp_bf16x4 *>(&act[r32][32 * mb + 8 * g + 4 * h]) = o[mb][g];
      mlp_wave_sync();  // the layer's input is in place
      const __bf16(*w)[kMlpRowH] = l == 0 ? m.w2 : (l == 1 ? m.w3 : m.w4);
      const float *b = l == 0 ? m.b2 : (l == 1 ? m.b3 : m.b4);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        mlp_act_block(mlp_block<8, kMlpRowH, true>(w, b, nullptr, act, mb, lane), o[mb]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // output layer: this lane's half of the units of its point, then the
    // other half from lane ^ 32
    float part = 0.0f;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int i = 0; i < 4; ++i) part += m.w5[32 * mb + 8 * g + 4 * h + i] * static_cast<float>(o[mb][g][i]);
    const float y = part + __shfl_xor(part, 32) + m.b5;
    if (h == nb) mine = y;
  }
  return mine;
}

}  // namespace vgm
