// Learned gating closure: a per-compartment MLP 8 -> 64 -> 64 -> 5 (tanh)
// that corrects the exponential-Euler gate update of the staggered step at
// large dt.  No counterpart in the reference.
//
// Features of one compartment at a coarse step n -> n+1 of length dt:
//   f = (V_{n+1}, Na p_n, Na q_n, KDR p_n, CaL p_n, CaL q_n, ln Ci_{n+1}, dt),
// normalised as (f - mu) / sd.  The network's 5 outputs c (times per-output
// scales) rescale the plain exponential-Euler increments of the 5 gates,
// delta = x_inf + (x - x_inf) e^{dt lam} - x at V_{n+1} (the staggered step's
// own update): x_{n+1} = x_n + delta (1 + c).  At a fixed point delta = 0,
// so the closure cannot drift a resting cell.
//
// Parameters (FP32 master copy, 5061 values): W1[8][64] b1[64] W2[64][64]
// b2[64] W3[64][5] b3[5], W[i][o] = weight from input i to output o.
//
// Kernels:
//   closure_data      training triples from fine trajectories (features, the
//                     plain increment delta, the fine end state minus the
//                     plain update)
//   closure_grad      FP32 forward + backward of the delta-weighted MSE loss
//                     sum_o (delta_o c_o - y_o)^2 on a minibatch, per-block
//                     partial gradients (deterministic)
//   closure_reduce    sums the partials; closure_adam: Adam step
//   closure_fwd32     FP32 inference (reference for the MFMA path)
//   closure_apply     BF16 MFMA inference with FP32 accumulation, fused with
//                     feature extraction and the plain gate update, in place
//                     after a staggered step that skipped its gates

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "bc_channels.h"
#include "bc_common.h"
#include "braincell_amd.h"

namespace {

constexpr int NIN = 8, NH = 64, NOUT = 5;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NOUT, NPARAM = OFF_B3 + NOUT;
static_assert(NPARAM == BC_CLOSURE_NPARAM, "parameter count");

// Gate slots of the Na/KDR/IL/CaL/pool state layout.
constexpr int S_NAP = 0, S_NAQ = 1, S_KDP = 2, S_CLP = 3, S_CLQ = 4, S_CI = 5, NSTATE = 6;

// The MFMA inference evaluates tanh z = 1 - 2 r with r = 1 / (1 + 2^{s z}),
// s = 2 log2(e), and folds both affine steps into the weights when they are
// loaded: a layer feeding a tanh has its weights and bias scaled by s (the
// MFMA returns s z), and the layer after it takes r instead of tanh z, with
// weights -2 W and bias b + sum_k W[k][.].  One hidden unit then costs an
// add, v_exp_f32 and v_rcp_f32; r saturates correctly at 0 and 1.
constexpr float kTanhScale = 2.8853900817779268f;  // 2 log2(e)
__device__ __forceinline__ float tanh_r(float sz) { return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(sz)); }

// Column sums of the layer-2 and layer-3 weights (the folded biases), into
// cs[0 .. NH + NOUT) of LDS, by the threads of a workgroup.
__device__ __forceinline__ void column_sums(const float* P, float* cs) {
  for (int t = threadIdx.x; t < NH + NOUT; t += blockDim.x) {
    float a = 0.0f;
    if (t < NH) {
      for (int k = 0; k < NH; ++k) a += P[OFF_W2 + k * NH + t];
    } else {
      for (int k = 0; k < NH; ++k) a += P[OFF_W3 + k * NOUT + (t - NH)];
    }
    cs[t] = a;
  }
}

// Plain exponential-Euler update of the 5 gates at voltage V: the staggered
// step's phase-C formulas under its FP32 math policy (hardware exponential).
__device__ __forceinline__ void plain_gates(const bc_chan_const<float>& cc, float V, const float* x, float dt,
                                            float* out) {
  float ap, bp, aq, bq;
  bc_na_rates<float, bc_fastmath<float>>(V, cc.na_vsh, ap, bp, aq, bq);
  out[0] = bc_gate_ab_cf<float, bc_fastmath<float>>(x[0], ap, bp, cc.na_phi, dt);
  out[1] = bc_gate_ab_cf<float, bc_fastmath<float>>(x[1], aq, bq, cc.na_phi, dt);
  float kp, kb;
  bc_kdr_rates<float, bc_fastmath<float>>(V, cc.kdr_vsh, kp, kb);
  out[2] = bc_gate_ab_cf<float, bc_fastmath<float>>(x[2], kp, kb, cc.kdr_phi, dt);
  float pi, pt, qi, qt;
  bc_cal_rates<float, bc_fastmath<float>>(V, cc.cal_vsh, pi, pt, qi, qt);
  out[3] = bc_gate_it_cf<float, bc_fastmath<float>>(x[3], pi, pt, cc.cal_phi_p, dt);
  out[4] = bc_gate_it_cf<float, bc_fastmath<float>>(x[4], qi, qt, cc.cal_phi_q, dt);
}

// ------------------------------------------------------------ training data
__global__ void closure_data(int64_t n, float dt, bc_chan_const<float> cc, const float* __restrict__ V1,
                             const float* __restrict__ S0, const float* __restrict__ S1, int64_t ld,
                             float* __restrict__ X, float* __restrict__ D, float* __restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x0[5], plain[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) x0[k] = S0[k * ld + i];
  const float V = V1[i];
  plain_gates(cc, V, x0, dt, plain);
  X[i * NIN + 0] = V;
#pragma unroll
  for (int k = 0; k < 5; ++k) X[i * NIN + 1 + k] = x0[k];
  X[i * NIN + 6] = logf(S1[S_CI * ld + i]);
  X[i * NIN + 7] = dt;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    D[i * NOUT + k] = plain[k] - x0[k];
    Y[i * NOUT + k] = S1[k * ld + i] - plain[k];
  }
}

// ------------------------------------------------------------ FP32 training
constexpr int TB = 128;    // samples (threads) per block
constexpr int HP = NH + 1; // padded LDS row

__global__ void __launch_bounds__(TB) closure_grad(int64_t B, const float* __restrict__ X,
                                                    const float* __restrict__ Dn, const float* __restrict__ Yn,
                                                    const int64_t* __restrict__ idx,
                                                    const float* __restrict__ P, float* __restrict__ partial,
                                                    float* __restrict__ loss_part) {
  __shared__ float W[NPARAM];
  __shared__ float xs[TB][NIN + 1];
  __shared__ float h1[TB][HP];
  __shared__ float h2[TB][HP];
  __shared__ float d3[TB][NOUT + 1];
  __shared__ float red[TB];
  const int t = threadIdx.x;
  for (int p = t; p < NPARAM; p += TB) W[p] = P[p];
  const int64_t s = (int64_t)blockIdx.x * TB + t;
  const bool valid = s < B;
  const int64_t row = valid ? idx[s] : 0;
#pragma unroll
  for (int i = 0; i < NIN; ++i) xs[t][i] = valid ? X[row * NIN + i] : 0.0f;
  __syncthreads();
  // forward
  for (int o = 0; o < NH; ++o) {
    float a = W[OFF_B1 + o];
#pragma unroll
    for (int i = 0; i < NIN; ++i) a += xs[t][i] * W[OFF_W1 + i * NH + o];
    h1[t][o] = tanhf(a);
  }
  for (int o = 0; o < NH; ++o) {
    float a = W[OFF_B2 + o];
    for (int i = 0; i < NH; ++i) a += h1[t][i] * W[OFF_W2 + i * NH + o];
    h2[t][o] = tanhf(a);
  }
  float loss = 0.0f;
  const float inv_b = 1.0f / (float)B;
  for (int o = 0; o < NOUT; ++o) {
    float a = W[OFF_B3 + o];
    for (int i = 0; i < NH; ++i) a += h2[t][i] * W[OFF_W3 + i * NOUT + o];
    const float w = valid ? Dn[row * NOUT + o] : 0.0f;
    const float d = valid ? w * a - Yn[row * NOUT + o] : 0.0f;
    loss += d * d;
    d3[t][o] = 2.0f * d * w * inv_b;
  }
  red[t] = loss * inv_b;
  __syncthreads();
  float* g = partial + (int64_t)blockIdx.x * NPARAM;
  // layer 3 gradients
  for (int p = t; p < NH * NOUT + NOUT; p += TB) {
    float a = 0.0f;
    if (p < NH * NOUT) {
      const int i = p / NOUT, o = p % NOUT;
      for (int k = 0; k < TB; ++k) a += h2[k][i] * d3[k][o];
      g[OFF_W3 + p] = a;
    } else {
      const int o = p - NH * NOUT;
      for (int k = 0; k < TB; ++k) a += d3[k][o];
      g[OFF_B3 + o] = a;
    }
  }
  if (t == 0) {
    float a = 0.0f;
    for (int k = 0; k < TB; ++k) a += red[k];
    loss_part[blockIdx.x] = a;
  }
  __syncthreads();
  // delta2 overwrites h2's row (element-wise in place)
  for (int i = 0; i < NH; ++i) {
    float a = 0.0f;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) a += W[OFF_W3 + i * NOUT + o] * d3[t][o];
    h2[t][i] = a * (1.0f - h2[t][i] * h2[t][i]);
  }
  __syncthreads();
  // layer 2 gradients
  for (int p = t; p < NH * NH + NH; p += TB) {
    float a = 0.0f;
    if (p < NH * NH) {
      const int i = p / NH, o = p % NH;
      for (int k = 0; k < TB; ++k) a += h1[k][i] * h2[k][o];
      g[OFF_W2 + p] = a;
    } else {
      const int o = p - NH * NH;
      for (int k = 0; k < TB; ++k) a += h2[k][o];
      g[OFF_B2 + o] = a;
    }
  }
  __syncthreads();
  // delta1 overwrites h1's row (element-wise in place)
  for (int i = 0; i < NH; ++i) {
    float a = 0.0f;
    for (int o = 0; o < NH; ++o) a += W[OFF_W2 + i * NH + o] * h2[t][o];
    h1[t][i] = a * (1.0f - h1[t][i] * h1[t][i]);
  }
  __syncthreads();
  // layer 1 gradients
  for (int p = t; p < NIN * NH + NH; p += TB) {
    float a = 0.0f;
    if (p < NIN * NH) {
      const int i = p / NH, o = p % NH;
      for (int k = 0; k < TB; ++k) a += xs[k][i] * h1[k][o];
      g[OFF_W1 + p] = a;
    } else {
      const int o = p - NIN * NH;
      for (int k = 0; k < TB; ++k) a += h1[k][o];
      g[OFF_B1 + o] = a;
    }
  }
}

__global__ void closure_reduce(int nblk, const float* __restrict__ partial, const float* __restrict__ loss_part,
                               float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) {
    float a = 0.0f;
    for (int b = 0; b < nblk; ++b) a += partial[(int64_t)b * NPARAM + p];
    grad[p] = a;
  } else if (p == NPARAM) {
    float a = 0.0f;
    for (int b = 0; b < nblk; ++b) a += loss_part[b];
    loss[0] = a;
  }
}

__global__ void closure_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                             const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1,
                             float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.0f - b1) * gp;
  const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp * c1) / (sqrtf(vp * c2) + eps);
}

// ----------------------------------------------------------- FP32 inference
__global__ void closure_fwd32(int64_t n, const float* __restrict__ X, const float* __restrict__ P,
                              const float* __restrict__ norm, float* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[NIN], h1[NH], h2[NH];
#pragma unroll
  for (int i = 0; i < NIN; ++i) x[i] = (X[s * NIN + i] - norm[i]) / norm[NIN + i];
  for (int o = 0; o < NH; ++o) {
    float a = P[OFF_B1 + o];
#pragma unroll
    for (int i = 0; i < NIN; ++i) a += x[i] * P[OFF_W1 + i * NH + o];
    h1[o] = tanhf(a);
  }
  for (int o = 0; o < NH; ++o) {
    float a = P[OFF_B2 + o];
    for (int i = 0; i < NH; ++i) a += h1[i] * P[OFF_W2 + i * NH + o];
    h2[o] = tanhf(a);
  }
  for (int o = 0; o < NOUT; ++o) {
    float a = P[OFF_B3 + o];
    for (int i = 0; i < NH; ++i) a += h2[i] * P[OFF_W3 + i * NOUT + o];
    out[s * NOUT + o] = a * norm[2 * NIN + o];
  }
}

// ----------------------------------------------------------- MFMA inference
// v_mfma_f32_16x16x32_bf16, lane l = 16 g + n (g = l / 16, n = l % 16):
//   A[i][k]: i = n, k = 8 g + j;  B[k][c]: k = 8 g + j, c = n;
//   D[r'][c]: r' = 4 g + r (r < 4), c = n.
// The layers are computed transposed, H^T = W^T X^T, so the 16 samples of a
// tile are the MFMA columns: lane (g, n) ends a layer holding hidden
// features 16 T + 4 g + r (T = output tile, r < 4) of sample n.  The next
// layer's B operand for K chunk c takes, in lane group g, the 8 features
// 16 (2c + j/4) + 4 g + j%4 that the lane already holds; the weights (A)
// are gathered in the same K order, so no activation moves between lanes.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int hid_feature(int c, int g, int j) { return 16 * (2 * c + j / 4) + 4 * g + (j & 3); }

struct MlpRegs {
  bf16x8 a1[4];     // layer 1, output tiles T = 0..3 (K = 8 inputs + bias + padding)
  bf16x8 a2[4][2];  // layer 2, output tile T, K chunk c
  bf16x8 a3[2];     // layer 3 (5 outputs padded to 16), K chunk c
  floatx4 b2[4];    // layer-2 bias in the accumulator layout
  floatx4 b3;
};

__device__ __forceinline__ void load_mlp(MlpRegs& m, const float* P, const float* cs, int lane) {
  const int g = lane >> 4, n = lane & 15;
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    const int o = 16 * T + n;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 8 * g + j;
      float w = 0.0f;
      if (k < NIN) w = P[OFF_W1 + k * NH + o];
      else if (k == NIN) w = P[OFF_B1 + o];
      m.a1[T][j] = (__bf16)(kTanhScale * w);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        m.a2[T][c][j] = (__bf16)(-2.0f * kTanhScale * P[OFF_W2 + hid_feature(c, g, j) * NH + o]);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o2 = 16 * T + 4 * g + r;
      m.b2[T][r] = kTanhScale * (P[OFF_B2 + o2] + cs[o2]);
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j)
      m.a3[c][j] = n < NOUT ? (__bf16)(-2.0f * P[OFF_W3 + hid_feature(c, g, j) * NOUT + n]) : (__bf16)0.0f;
#pragma unroll
  for (int r = 0; r < 4; ++r) m.b3[r] = (4 * g + r < NOUT) ? P[OFF_B3 + 4 * g + r] + cs[NH + 4 * g + r] : 0.0f;
}

// One tile of 16 samples.  xin: in lane group 0, the normalised features of
// sample n; returns D3 (lane (g, n) holds outputs 4 g + r of sample n).
__device__ __forceinline__ floatx4 mlp_tile(const MlpRegs& m, int lane, const float* xin) {
  const int g = lane >> 4;
  bf16x8 b;
#pragma unroll
  for (int j = 0; j < 8; ++j) b[j] = (__bf16)(g == 0 ? xin[j] : (g == 1 && j == 0 ? 1.0f : 0.0f));
  floatx4 h[4];
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.a1[T], b, z, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) h[T][r] = tanh_r(z[r]);
  }
  bf16x8 hb[2];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j) hb[c][j] = (__bf16)h[2 * c + j / 4][j & 3];
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    floatx4 z = m.b2[T];
    z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.a2[T][0], hb[0], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.a2[T][1], hb[1], z, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) h[T][r] = tanh_r(z[r]);
  }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j) hb[c][j] = (__bf16)h[2 * c + j / 4][j & 3];
  floatx4 y = m.b3;
  y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.a3[0], hb[0], y, 0, 0, 0);
  y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.a3[1], hb[1], y, 0, 0, 0);
  return y;
}

// MFMA inference on precomputed raw features (test / measurement path).
// The weights are staged once per workgroup in LDS and gathered from there
// into each wave's operand registers.
__global__ void __launch_bounds__(256) closure_fwd_mfma(int64_t n, const float* __restrict__ X,
                                                        const float* __restrict__ P, const float* __restrict__ norm,
                                                        float* __restrict__ out) {
  __shared__ float Ps[NPARAM];
  __shared__ float cs[NH + NOUT];
  for (int p = threadIdx.x; p < NPARAM; p += blockDim.x) Ps[p] = P[p];
  __syncthreads();
  column_sums(Ps, cs);
  __syncthreads();
  MlpRegs m;
  const int lane = threadIdx.x & 63;
  load_mlp(m, Ps, cs, lane);
  const int g = lane >> 4, cn = lane & 15;
  const int64_t n_tiles = (n + 15) / 16;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t tile = wave; tile < n_tiles; tile += n_waves) {
    const int64_t s = tile * 16 + cn;
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = (g == 0 && s < n) ? (X[s * NIN + i] - norm[i]) / norm[NIN + i] : 0.0f;
    const floatx4 y = mlp_tile(m, lane, x);
    if (s < n) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = 4 * g + r;
        if (o < NOUT) out[s * NOUT + o] = y[r] * norm[2 * NIN + o];
      }
    }
  }
}

// Closure gate update, in place, after a staggered step run with
// BC_STEP_SKIP_GATES: S holds the step-start gates and the updated Ci, V1
// the new voltage.  A wave takes 64 compartments at a time (four MFMA
// tiles): every lane extracts the features and the plain update of its own
// compartment (coalesced loads, no idle lanes), the features move to the
// operand lanes of their tile by shuffles, the four tiles run the network,
// and the outputs shuffle back.  The gate becomes clip(x + delta (1 + c),
// 0, 1) with delta the plain increment.
__global__ void __launch_bounds__(256) closure_apply(int64_t n, float dt, bc_chan_const<float> cc,
                                                     const float* __restrict__ V1, float* __restrict__ S, int64_t ld,
                                                     const float* __restrict__ P, const float* __restrict__ norm,
                                                     uint8_t* __restrict__ flags, float flag_tol) {
  __shared__ float Ps[NPARAM];
  __shared__ float cs[NH + NOUT];
  for (int p = threadIdx.x; p < NPARAM; p += blockDim.x) Ps[p] = P[p];
  __syncthreads();
  column_sums(Ps, cs);
  __syncthreads();
  MlpRegs m;
  const int lane = threadIdx.x & 63;
  load_mlp(m, Ps, cs, lane);
  const int g = lane >> 4, cn = lane & 15;
  float mu[8], isd[8], sc[NOUT];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    mu[i] = norm[i];
    isd[i] = 1.0f / norm[NIN + i];
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o) sc[o] = norm[2 * NIN + o];
  const int64_t n_quads = (n + 63) / 64;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t quad = wave; quad < n_quads; quad += n_waves) {
    const int64_t s = quad * 64 + lane;
    const bool ok = s < n;
    float x0[NOUT], plain[NOUT], f[8];
    if (ok) {
#pragma unroll
      for (int k = 0; k < NOUT; ++k) x0[k] = S[k * ld + s];
      const float V = V1[s];
      plain_gates(cc, V, x0, dt, plain);
      f[0] = V;
#pragma unroll
      for (int k = 0; k < NOUT; ++k) f[1 + k] = x0[k];
      f[6] = __logf(S[S_CI * ld + s]);
      f[7] = dt;
#pragma unroll
      for (int i = 0; i < 8; ++i) f[i] = (f[i] - mu[i]) * isd[i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) f[i] = 0.0f;
#pragma unroll
      for (int k = 0; k < NOUT; ++k) x0[k] = plain[k] = 0.0f;
    }
    float c[NOUT];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float xq[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) xq[i] = __shfl(f[i], 16 * q + cn);
      const floatx4 y = mlp_tile(m, lane, xq);
      float t[NOUT];
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = __shfl(y[k], cn);
      t[4] = __shfl(y[0], 16 + cn);
      if (g == q) {
#pragma unroll
        for (int k = 0; k < NOUT; ++k) c[k] = t[k];
      }
    }
    if (ok) {
      bool over = false;
#pragma unroll
      for (int k = 0; k < NOUT; ++k) {
        const float raw = plain[k] + (plain[k] - x0[k]) * (c[k] * sc[k]);
        over |= !(raw >= -flag_tol && raw <= 1.0f + flag_tol);  // also true for NaN
        S[k * ld + s] = fminf(fmaxf(raw, 0.0f), 1.0f);
      }
      if (over && flags != nullptr) flags[s] = 1;
    }
  }
}

int grid_for(int64_t n, int per) {
  int64_t b = (n + per - 1) / per;
  return (int)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

// Scratch of bc_closure_grad_noalloc: the per-block partial gradients and
// losses of closure_grad in a file-scope device array, so that the entry
// point needs no workspace from its caller.  Each GPU holds its own copy for
// the life of the process; every call on a GPU reuses it.
constexpr int kScratchBlocks = BC_CLOSURE_NOALLOC_MAX_BATCH / TB;
static_assert(kScratchBlocks * TB == BC_CLOSURE_NOALLOC_MAX_BATCH, "batch bound is a whole number of blocks");
__device__ __align__(256) float scratch[(size_t)kScratchBlocks * NPARAM + kScratchBlocks];

// Device address of `scratch` on the current device, resolved once per device
// with hipGetSymbolAddress (the array's name is a host-side stand-in in host
// code and must never be passed to a kernel or a copy).
constexpr int kMaxDevices = 64;
float* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_address(float** out) {
  int dev = -1;
  BC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return bc::fail(BC_UNSUPPORTED, "device ordinal %d >= %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    BC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) return bc::fail(BC_HIP_ERROR, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[dev];
  return BC_OK;
}

int closure_grad_launch(const bc_closure_train_desc* d, int nblk, float* partial, float* loss_part,
                        hipStream_t st) {
  hipLaunchKernelGGL(closure_grad, dim3(nblk), dim3(TB), 0, st, d->batch, static_cast<const float*>(d->X),
                     static_cast<const float*>(d->Dn), static_cast<const float*>(d->Yn), d->idx,
                     static_cast<const float*>(d->P), partial, loss_part);
  BC_HIP(hipGetLastError());
  hipLaunchKernelGGL(closure_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, st, nblk, partial, loss_part,
                     static_cast<float*>(d->grad), static_cast<float*>(d->loss));
  BC_HIP(hipGetLastError());
  return bc::ok();
}

}  // namespace

extern "C" {

int bc_closure_data(const bc_closure_data_desc* d, void* stream) {
  if (d == nullptr || d->chan_const == nullptr) return bc::fail(BC_INVALID_ARG, "desc / chan_const is NULL");
  if (d->n < 0 || d->ld < d->n) return bc::fail(BC_INVALID_ARG, "need 0 <= n <= ld");
  if (d->n == 0) return bc::ok();
  if (!d->V1 || !d->S0 || !d->S1 || !d->X || !d->D || !d->Y)
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  const int64_t nb = (d->n + 255) / 256;
  hipLaunchKernelGGL(closure_data, dim3((unsigned)nb), dim3(256), 0, static_cast<hipStream_t>(stream), d->n,
                     (float)d->dt, *static_cast<const bc_chan_const<float>*>(d->chan_const),
                     static_cast<const float*>(d->V1), static_cast<const float*>(d->S0),
                     static_cast<const float*>(d->S1), d->ld, static_cast<float*>(d->X), static_cast<float*>(d->D),
                     static_cast<float*>(d->Y));
  BC_HIP(hipGetLastError());
  return bc::ok();
}

int bc_closure_grad(const bc_closure_train_desc* d, void* stream) {
  if (d == nullptr) return bc::fail(BC_INVALID_ARG, "desc is NULL");
  if (d->batch <= 0) return bc::fail(BC_INVALID_ARG, "batch must be > 0");
  if (!d->X || !d->Dn || !d->Yn || !d->idx || !d->P || !d->partial || !d->loss_part || !d->grad || !d->loss)
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  const int nblk = (int)((d->batch + TB - 1) / TB);
  if (nblk > d->max_blocks) return bc::fail(BC_INVALID_ARG, "batch needs %d partial blocks > max_blocks %d", nblk,
                                            d->max_blocks);
  return closure_grad_launch(d, nblk, static_cast<float*>(d->partial), static_cast<float*>(d->loss_part),
                             static_cast<hipStream_t>(stream));
}

int bc_closure_grad_noalloc(const bc_closure_train_desc* d, void* stream) {
  if (d == nullptr) return bc::fail(BC_INVALID_ARG, "desc is NULL");
  if (d->batch <= 0) return bc::fail(BC_INVALID_ARG, "batch must be > 0");
  if (d->batch > BC_CLOSURE_NOALLOC_MAX_BATCH)
    return bc::fail(BC_INVALID_ARG, "batch %lld exceeds the %d samples the scratch array holds",
                    (long long)d->batch, BC_CLOSURE_NOALLOC_MAX_BATCH);
  if (!d->X || !d->Dn || !d->Yn || !d->idx || !d->P || !d->grad || !d->loss)
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  float* base = nullptr;
  int rc = scratch_address(&base);
  if (rc != BC_OK) return rc;
  const int nblk = (int)((d->batch + TB - 1) / TB);
  // closure_grad writes every partial of its nblk blocks before closure_reduce
  // reads them, so nothing in the scratch needs clearing between calls
  return closure_grad_launch(d, nblk, base, base + (size_t)kScratchBlocks * NPARAM, static_cast<hipStream_t>(stream));
}

int bc_closure_adam(const bc_closure_adam_desc* d, void* stream) {
  if (d == nullptr || !d->P || !d->m || !d->v || !d->grad) return bc::fail(BC_INVALID_ARG, "NULL argument");
  if (d->step < 1) return bc::fail(BC_INVALID_ARG, "step counts from 1");
  const double c1 = 1.0 / (1.0 - pow(d->beta1, (double)d->step));
  const double c2 = 1.0 / (1.0 - pow(d->beta2, (double)d->step));
  hipLaunchKernelGGL(closure_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<float*>(d->P), static_cast<float*>(d->m), static_cast<float*>(d->v),
                     static_cast<const float*>(d->grad), (float)d->lr, (float)d->beta1, (float)d->beta2,
                     (float)d->eps, (float)c1, (float)c2);
  BC_HIP(hipGetLastError());
  return bc::ok();
}

int bc_closure_forward(const bc_closure_fwd_desc* d, void* stream) {
  if (d == nullptr || !d->X || !d->P || !d->norm || !d->out) return bc::fail(BC_INVALID_ARG, "NULL argument");
  if (d->n < 0) return bc::fail(BC_INVALID_ARG, "n < 0");
  if (d->n == 0) return bc::ok();
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->mfma) {
    hipLaunchKernelGGL(closure_fwd_mfma, dim3(grid_for((d->n + 15) / 16, 128)), dim3(256), 0, st, d->n,
                       static_cast<const float*>(d->X), static_cast<const float*>(d->P),
                       static_cast<const float*>(d->norm), static_cast<float*>(d->out));
  } else {
    hipLaunchKernelGGL(closure_fwd32, dim3(grid_for(d->n, 256)), dim3(256), 0, st, d->n,
                       static_cast<const float*>(d->X), static_cast<const float*>(d->P),
                       static_cast<const float*>(d->norm), static_cast<float*>(d->out));
  }
  BC_HIP(hipGetLastError());
  return bc::ok();
}

int bc_closure_apply(const bc_closure_apply_desc* d, void* stream) {
  if (d == nullptr || !d->V1 || !d->S || !d->P || !d->norm || !d->chan_const)
    return bc::fail(BC_INVALID_ARG, "NULL argument");
  if (d->n < 0 || d->ld < d->n) return bc::fail(BC_INVALID_ARG, "need 0 <= n <= ld");
  if (d->n == 0) return bc::ok();
  hipLaunchKernelGGL(closure_apply, dim3(grid_for((d->n + 63) / 64, 32)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), d->n, (float)d->dt,
                     *static_cast<const bc_chan_const<float>*>(d->chan_const), static_cast<const float*>(d->V1),
                     static_cast<float*>(d->S), d->ld, static_cast<const float*>(d->P),
                     static_cast<const float*>(d->norm), static_cast<uint8_t*>(d->flags), (float)d->flag_tol);
  BC_HIP(hipGetLastError());
  return bc::ok();
}

}  // extern "C"
