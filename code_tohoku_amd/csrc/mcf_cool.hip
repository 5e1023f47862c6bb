// mcf_cool.hip -- training and evaluation kernels of the KI-cooling
// surrogate (mcf_mlp.h, mcf_cool.h).
//
//   k_cool_samples  training samples from the reference's IMC integrator
//   k_mlp_pack      FP32 master weights -> BF16 MFMA operands
//   k_mlp_infer     forward pass over a batch (MFMA)
//   k_mlp_grad      loss and gradient of a batch: MFMA forward, FP32 backward
//   k_mlp_adam      Adam step on the master weights
//   k_cool_solve    per-cell backward-Euler cooling solves, IMC or surrogate
#include "mcf_cool.h"
#include "mcf_kernels.h"

namespace mcf {

// counter-based uniform in [0,1) (splitmix64 finaliser of seed and index)
__device__ __forceinline__ double u01(unsigned long long seed, unsigned long long idx) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (idx + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Sample i: log-uniform N in [0.1, 1e4], T in [10, 1e4] K, dt/t_cool in
// [1e-3, 1e3]; gas of H only (gammi1 = 2/3, n = N).  X[i] = scaled features,
// Y[i] = scaled log10(T_end/T) of IMC.
__global__ void k_cool_samples(long n, unsigned long long seed, float *X, float *Y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lN = -1.0 + 5.0 * u01(seed, 3 * i);
  const double lT = 1.0 + 3.0 * u01(seed, 3 * i + 1);
  const double lr = -3.0 + 6.0 * u01(seed, 3 * i + 2);
  const double N = exp10(lN), T0 = exp10(lT) * 1.0e-3, r = exp10(lr);
  const double rho = 1.27 * N, nn = N, gi = 2.0 / 3.0;
  const double P0 = KB * nn * T0;
  const double L0 = fcool_ki(rho, T0);
  double dt = r * P0 / (gi * fabs(L0));
  if (!(dt < 1.0e300)) dt = 0.0;
  const double P = imc(P0, nn, dt, rho, gi);
  float u0, u1, u2;
  sur_features(P0, nn, dt, rho, gi, u0, u1, u2);
  X[3 * i] = u0;
  X[3 * i + 1] = u1;
  X[3 * i + 2] = u2;
  Y[i] = (float)(log10(P / P0) / (double)MLP_Y_S);
}

__global__ void k_mlp_pack(const float *w, __bf16 *pk) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_NPK) return;
  int j = e & 7, l = (e >> 3) & 63;
  float v;
  if (e < MLP_PK2) {
    int hb = e >> 9;
    v = mlp_pk1_val(w, hb, l, j);
  } else {
    int r = (e - MLP_PK2) >> 9; // ob * 4 + kb
    v = mlp_pk2_val(w, r >> 2, r & 3, l, j);
  }
  pk[e] = mlp_bf16(v);
}

// One wave per 64 samples; lanes past n evaluate zeros and store nothing.
__global__ __launch_bounds__(256) void k_mlp_infer(const __bf16 *pk, const float *w, const float *X, long n, float *Y) {
  long base = ((long)blockIdx.x * blockDim.x + threadIdx.x) & ~63L;
  long i = base + (threadIdx.x & 63);
  bool ok = i < n;
  float u0 = ok ? X[3 * i] : 0.0f, u1 = ok ? X[3 * i + 1] : 0.0f, u2 = ok ? X[3 * i + 2] : 0.0f;
  float y = mlp_wave_forward(pk, w, u0, u1, u2);
  if (ok) Y[i] = y;
}

// Mean-squared-error gradient over samples [0, n) with weight inv_b = 1/B:
// each 64-thread block takes chunks of 64 samples.  The forward pass is the
// inference pass (BF16 operands, FP32 accumulation); the backward pass is
// FP32 and treats the BF16 roundings as identity (straight-through), with
// the rounded W1, b1, W2 and h1 the forward pass used.
constexpr int GLD = 65; // LDS pitch
__global__ __launch_bounds__(64) void k_mlp_grad(const __bf16 *pk, const float *w, const float *X,
                                                   const float *Yt, long n, float inv_b, float *grad,
                                                   float *loss) {
  __shared__ float z1[64 * GLD], z2[64 * GLD], hh[64 * GLD], dz[64 * GLD];
  __shared__ float xs[64][4], dys[64];
  __shared__ float w2b[64 * 64];
  const int t = threadIdx.x;
  for (int e = t; e < 4096; e += 64) w2b[e] = (float)mlp_bf16(w[MLP_OW2 + e]);
  float gW2[64];
#pragma unroll
  for (int e = 0; e < 64; ++e) gW2[e] = 0.0f;
  float gW3 = 0.0f, gb2 = 0.0f, gb1 = 0.0f, gW1[3] = {0.0f, 0.0f, 0.0f}, gb3 = 0.0f, ls = 0.0f;
  __syncthreads();
  for (long base = 64L * blockIdx.x; base < n; base += 64L * gridDim.x) {
    const long i = base + t;
    const bool ok = i < n;
    float u0 = ok ? X[3 * i] : 0.0f, u1 = ok ? X[3 * i + 1] : 0.0f, u2 = ok ? X[3 * i + 2] : 0.0f;
    float y = mlp_wave_forward(pk, w, u0, u1, u2, z1, z2, GLD);
    float dy = ok ? 2.0f * (y - Yt[i]) * inv_b : 0.0f;
    if (ok) ls += (y - Yt[i]) * (y - Yt[i]) * inv_b;
    xs[t][0] = (float)mlp_bf16(u0); xs[t][1] = (float)mlp_bf16(u1); xs[t][2] = (float)mlp_bf16(u2); xs[t][3] = 1.0f;
    dys[t] = dy;
    __syncthreads();
    // per sample t: h1 = bf16(silu(z1)), dz2 = W3 * dy * silu'(z2)
    for (int o = 0; o < 64; ++o) {
      float a = z1[t * GLD + o];
      hh[t * GLD + o] = (float)mlp_bf16(mlp_silu(a));
      float b = z2[t * GLD + o];
      float sg = mlp_sigmoid(b);
      dz[t * GLD + o] = w[MLP_OW3 + o] * dy * (sg * (1.0f + b * (1.0f - sg)));
    }
    __syncthreads();
    // per output unit t: dW3, db2, dW2 row t
    for (int s = 0; s < 64; ++s) {
      float d = dz[s * GLD + t];
      gW3 += dys[s] * mlp_silu(z2[s * GLD + t]);
      gb2 += d;
#pragma unroll
      for (int e = 0; e < 64; ++e) gW2[e] += d * hh[s * GLD + e];
    }
    gb3 += dy;
    __syncthreads();
    // per sample t: dh1 = W2b^T dz2, dz1 = dh1 * silu'(z1) -> z2 (free now)
    for (int e = 0; e < 64; ++e) {
      float acc = 0.0f;
      for (int o = 0; o < 64; ++o) acc += w2b[o * 64 + e] * dz[t * GLD + o];
      float a = z1[t * GLD + e];
      float sg = mlp_sigmoid(a);
      z2[t * GLD + e] = acc * (sg * (1.0f + a * (1.0f - sg)));
    }
    __syncthreads();
    // per hidden unit t: db1, dW1 row t
    for (int s = 0; s < 64; ++s) {
      float d = z2[s * GLD + t];
      gb1 += d;
      gW1[0] += d * xs[s][0];
      gW1[1] += d * xs[s][1];
      gW1[2] += d * xs[s][2];
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < 64; ++e) atomicAdd(&grad[MLP_OW2 + 64 * t + e], gW2[e]);
  atomicAdd(&grad[MLP_OW3 + t], gW3);
  atomicAdd(&grad[MLP_OB2 + t], gb2);
  atomicAdd(&grad[MLP_OB1 + t], gb1);
  for (int k = 0; k < 3; ++k) atomicAdd(&grad[MLP_OW1 + 3 * t + k], gW1[k]);
  // db3 and the loss: wave sums
  for (int off = 32; off > 0; off >>= 1) {
    gb3 += __shfl_xor(gb3, off, 64);
    ls += __shfl_xor(ls, off, 64);
  }
  if (t == 0) {
    atomicAdd(&grad[MLP_OB3], gb3);
    atomicAdd(loss, ls);
  }
}

__global__ void k_mlp_adam(float *w, float *m, float *v, const float *g, float lr, float b1t, float b2t) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_NP) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1.0e-8f;
  float ge = g[e];
  float me = b1 * m[e] + (1.0f - b1) * ge;
  float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  float mh = me / (1.0f - b1t), vh = ve / (1.0f - b2t);
  w[e] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_fill_f32(float *a, int n, float v) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = v;
}

// Backward-Euler cooling solves P + gammi1*Lambda(P)*dt = P0 for n cells.
// mode 0: IMC.  mode 1: surrogate + Newton + residual test; rejected cells
// (flag[i] = 1) are queued in fq for k_cool_fallback, which runs IMC on
// dense waves.
__global__ __launch_bounds__(256) void k_cool_solve(const double *P0, const double *nn, const double *dt,
                                                    const double *rho, const double *gi, long n, int mode,
                                                    const __bf16 *pk, const float *w, double *Pout, int *flag,
                                                    unsigned *nfb, int *fq) {
  long base = ((long)blockIdx.x * blockDim.x + threadIdx.x) & ~63L;
  const int lane = threadIdx.x & 63;
  long i = base + lane;
  bool ok = i < n;
  double p0 = ok ? P0[i] : 1.0, n0 = ok ? nn[i] : 1.0, d0 = ok ? dt[i] : 0.0, r0 = ok ? rho[i] : 1.27;
  double g0 = ok ? gi[i] : 2.0 / 3.0;
  if (mode == 0) {
    if (ok) {
      Pout[i] = imc(p0, n0, d0, r0, g0);
      if (flag) flag[i] = 0;
    }
    return;
  }
  float u0, u1, u2;
  sur_features(p0, n0, d0, r0, g0, u0, u1, u2);
  if (!ok) u0 = u1 = u2 = 0.0f;
  float y = mlp_wave_forward(pk, w, u0, u1, u2);
  double P;
  const bool fell = !sur_refine(p0, n0, d0, r0, g0, y, P);
  const unsigned long long mf = __ballot(ok && fell);
  if (mf) {
    const int leader = __ffsll((long long)mf) - 1;
    unsigned b = 0;
    if (lane == leader) b = atomicAdd(nfb, (unsigned)__popcll(mf));
    b = __shfl(b, leader, 64);
    if (ok && fell) fq[b + (unsigned)__popcll(mf & ((1ull << lane) - 1ull))] = (int)i;
  }
  if (ok) {
    if (!fell) Pout[i] = P;
    if (flag) flag[i] = fell ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void k_cool_fallback(const double *P0, const double *nn, const double *dt,
                                                       const double *rho, const double *gi, const unsigned *nfb,
                                                       const int *fq, double *Pout) {
  const unsigned nf = *nfb;
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < nf; e += stride) {
    const int i = fq[e];
    Pout[i] = imc(P0[i], nn[i], dt[i], rho[i], gi[i]);
  }
}

} // namespace mcf
