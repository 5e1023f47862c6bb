// mcf_mlp.h -- the KI-cooling surrogate: an MLP 3 -> 64 -> 64 -> 1 with SiLU,
// evaluated by one wavefront for 64 cells at a time on the gfx950 matrix
// cores (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32 accumulation).
//
// Inputs (per cell): u0 = (log10 N - 1.5)/2.5, u1 = (log10 T[K] - 2.5)/1.5,
// u2 = log10(dt/t_cool)/3, with N = rho/1.27 the density of the KI cooling
// function, T the temperature before the update, dt the step of the
// backward-Euler solve and t_cool = P/(gammi1*|Lambda(N,T)|).
// Output: log10(T_end/T)/3, T_end the solution of the backward-Euler
// equation P_end + gammi1*Lambda(N, T_end)*dt = P solved by IMC
// (3DMHD-CfCA/MCF3DSGincfca3.f90:2898-2921).
//
// Everything is computed transposed (hidden units as matrix rows, cells as
// columns), so the accumulator layout of one layer is directly the B
// operand of the next:
//   layer 1: Z1 = [W1 | b1] [u; 1]       A = weights (32 x 16, 4 of K used)
//   layer 2: Z2 = W2 silu(Z1) + b2       A = W2 with its columns permuted
//   layer 3: y  = W3 . silu(Z2) + b3     VALU dot product, one lane exchange
// MFMA lane layout (32x32x16): lane l holds A row l%32 and B column l%32
// with 8 K values each in lane half l/32; D element i of lane l is row
// 8*(i/4) + 4*(l/32) + i%4, column l%32.  The K order inside one
// instruction only has to agree between A and B, so the 16 hidden units of
// K block kb are taken in the order accumulator values leave layer 1:
// element j of lane half h is hidden unit 16*kb + 8*(j/4) + 4*h + j%4.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcf {

constexpr int MLP_H = 64;
// FP32 master parameters, flat: W1[64][3] b1[64] W2[64][64] b2[64] W3[64] b3
constexpr int MLP_OW1 = 0, MLP_OB1 = 192, MLP_OW2 = 256, MLP_OB2 = 4352, MLP_OW3 = 4416, MLP_OB3 = 4480;
constexpr int MLP_NP = 4481;
// packed BF16 A operands: layer 1 [hb 2][lane 64][8], layer 2 [ob 2][kb 4][lane 64][8]
constexpr int MLP_PK1 = 0, MLP_PK2 = 2 * 64 * 8;
constexpr int MLP_NPK = MLP_PK2 + 2 * 4 * 64 * 8;

// feature / target scaling
constexpr float MLP_LN_C = 1.5f, MLP_LN_S = 2.5f;   // log10 N
constexpr float MLP_LT_C = 2.5f, MLP_LT_S = 1.5f;   // log10 T[K]
constexpr float MLP_LR_S = 3.0f;                    // log10 dt/t_cool
constexpr float MLP_Y_S = 3.0f;                     // log10 T_end/T

typedef float mlp_f16v __attribute__((ext_vector_type(16)));
typedef __bf16 mlp_b8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int mlp_perm(int kb, int h, int j) { return 16 * kb + 8 * (j >> 2) + 4 * h + (j & 3); }
// hidden row held by accumulator element i of lane half h in tile block hb
__device__ __forceinline__ int mlp_row(int hb, int h, int i) { return 32 * hb + 8 * (i >> 2) + 4 * h + (i & 3); }

__device__ __forceinline__ float mlp_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float mlp_silu(float z) { return z * mlp_sigmoid(z); }

// Round-to-nearest-even FP32 -> BF16 (the conversion the operands get).
__device__ __forceinline__ __bf16 mlp_bf16(float f) { return (__bf16)f; }

// Packed operand element (lane l, element j) of layer 1, hidden block hb.
__device__ __forceinline__ float mlp_pk1_val(const float *w, int hb, int l, int j) {
  const int h = l >> 5, m = 32 * hb + (l & 31), k = 8 * h + j;
  if (k < 3) return w[MLP_OW1 + 3 * m + k];
  if (k == 3) return w[MLP_OB1 + m];
  return 0.0f;
}
// Packed operand element of layer 2, output block ob, K block kb.
__device__ __forceinline__ float mlp_pk2_val(const float *w, int ob, int kb, int l, int j) {
  const int m = 32 * ob + (l & 31);
  return w[MLP_OW2 + 64 * m + mlp_perm(kb, l >> 5, j)];
}

struct MlpFwd {
  float y;          // output of this lane's cell
};

// Forward pass of the 64 cells of a wave; u0..u2 are this lane's inputs.
// Must be called by all 64 lanes with the same arguments for pk / w.
// If z1 / z2 are not null the pre-activations are stored as
// z[cell * ldz + hidden] (cell = lane of the wave) for the backward pass.
__device__ __forceinline__ float mlp_wave_forward(const __bf16 *__restrict__ pk, const float *__restrict__ w,
                                                  float u0, float u1, float u2, float *z1 = nullptr,
                                                  float *z2 = nullptr, int ldz = 0) {
  const int l = threadIdx.x & 63, h = l >> 5;
  const mlp_b8v *pk1 = reinterpret_cast<const mlp_b8v *>(pk + MLP_PK1);
  const mlp_b8v *pk2 = reinterpret_cast<const mlp_b8v *>(pk + MLP_PK2);
  float yown = 0.0f;
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int src = 32 * cb + (l & 31);
    const float v0 = __shfl(u0, src, 64), v1 = __shfl(u1, src, 64), v2 = __shfl(u2, src, 64);
    mlp_b8v b1;
#pragma unroll
    for (int j = 0; j < 8; ++j) b1[j] = mlp_bf16(0.0f);
    if (h == 0) { b1[0] = mlp_bf16(v0); b1[1] = mlp_bf16(v1); b1[2] = mlp_bf16(v2); b1[3] = mlp_bf16(1.0f); }
    mlp_f16v d1[2];
#pragma unroll
    for (int hb = 0; hb < 2; ++hb) {
      mlp_f16v z = {};
      d1[hb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pk1[hb * 64 + l], b1, z, 0, 0, 0);
    }
    if (z1) {
#pragma unroll
      for (int hb = 0; hb < 2; ++hb)
#pragma unroll
        for (int i = 0; i < 16; ++i) z1[src * ldz + mlp_row(hb, h, i)] = d1[hb][i];
    }
    mlp_b8v b2[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int j = 0; j < 8; ++j) b2[kb][j] = mlp_bf16(mlp_silu(d1[kb >> 1][8 * (kb & 1) + j]));
    float part = 0.0f;
#pragma unroll
    for (int ob = 0; ob < 2; ++ob) {
      mlp_f16v acc;
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = w[MLP_OB2 + mlp_row(ob, h, i)];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pk2[(ob * 4 + kb) * 64 + l], b2[kb], acc, 0, 0, 0);
      if (z2) {
#pragma unroll
        for (int i = 0; i < 16; ++i) z2[src * ldz + mlp_row(ob, h, i)] = acc[i];
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) part += w[MLP_OW3 + mlp_row(ob, h, i)] * mlp_silu(acc[i]);
    }
    part += __shfl_xor(part, 32, 64);
    const float y = part + w[MLP_OB3];
    if (h == cb) yown = y;
  }
  return yown;
}

} // namespace mcf
