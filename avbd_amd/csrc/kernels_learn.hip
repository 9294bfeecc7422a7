// kernels_learn.hip — learned warm start of contact lambdas on gfx950.
//
// A small MLP predicts the starting dual state (lambda_n, lambda_t,
// lambda_b, log10 k_n) of each box-box and floor contact group from 24
// local features. It only replaces the initial dual state written at emission;
// the iterations that follow are the plain solver's, so the converged answer
// is unchanged. No counterpart in the reference (its warm start is the
// persistent-lambda cache, kernels_6dof.py:2695-2912, which stays the
// fallback).
//
//  k_contact_features   24 features per dynamic group (below)
//  k_mlp_infer          one wavefront per 16 contacts: three layers as
//                       v_mfma_f32_16x16x32_bf16 (bf16 operands, fp32
//                       accumulation), weights staged in LDS, SiLU in fp32
//  k_mlp_train_tile     forward + backward of 64 samples per workgroup on
//                       v_mfma_f32_16x16x32_bf16 (bf16 weights, layer
//                       inputs and gradients, fp32 accumulation); each
//                       workgroup stores its gradient row, and
//  k_mlp_part_reduce    sums the rows in a fixed order
//  k_adam               Adam step on the fp32 master weights
//
// Network: x (24) -> 64 SiLU -> 64 SiLU -> 4. Parameters (fp32, in order):
// W1 [24][64], b1 [64], W2 [64][64], b2 [64], W3 [64][4], b3 [4], then the
// normalisation x_mean [24], x_std [24], y_mean [4], y_std [4].
//
// Features of a box-box group (bodies a = reference, b = incident;
// r_a = R_a off_a) or a floor group (body a, the floor as b):
//   0-2  contact normal n (world; +y on the floor)
//   3-5  off_a (body a frame)          6-8  off_b (body b frame; 0 on the floor)
//   9    C0 = n . (w_a - w_b) (on the floor: corner height above it)
//   10-12 v_rel along the rows' directions (n, t, b: the Duff basis of n for
//        box-box, world x and z on the floor); v_rel = velocity of the
//        contact point on a minus on b
//   13   1 / m_a     14  1 / m_b (0 on the floor)
//   15-17 slog(lambda n, t, b at the start of the substep)
//   18   contact age (substeps it has been cached; 0 = new or floor)
//   19-21 log10 k n, t, b            22 friction mu
//   23   kind: 0 box-box, 1 floor, -1 pin (pins are left to the plain start)
// Outputs are corrections to the plain warm start (the cached or persistent
// state the prelude decays), so that a zero output reproduces it exactly:
// with ag = alpha * gamma (the decay the prelude applies),
//   y_r = slog(lambda_r at the end of the substep) - slog(ag * cached lambda_r)
//   y_3 = log10(k_n at the end of the substep) - log10(clamp(gamma * cached k_n))
// slog(x) = sign(x) log(1 + |x|). Tangent outputs count only on contacts with
// friction rows (mask 0b1111, else 0b1001).
#include "avbd_internal.h"

namespace avbd {

__device__ __forceinline__ v3 f3(float4 a) { return mk3(a.x, a.y, a.z); }
__device__ __forceinline__ q4 f4q(float4 a) { return mk4(a.x, a.y, a.z, a.w); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ float slog(float x) {
  return x < 0.0f ? -log1pf(-x) : log1pf(x);
}
__device__ __forceinline__ float islog(float y) { return y < 0.0f ? -expm1f(-y) : expm1f(y); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
// round to bfloat16 (nearest even), kept in fp32
__device__ __forceinline__ float bf16r(float x) { return (float)(__bf16)x; }

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_contact_features(int n_groups, int g_first, const GroupGeo* __restrict__ geo,
                   const GroupSt* __restrict__ gst, const float4* __restrict__ pose,
                   const float4* __restrict__ vel, const float4* __restrict__ bstat,
                   float* __restrict__ feat) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_groups) return;
  const GroupGeo G = geo[g_first + e];
  const GroupSt S = gst[g_first + e];
  const int kind = g_kind(G.hdr.x);
  float* f = feat + (size_t)e * 24;
  if (kind == kGPin) {  // not predicted: all features zero, kind -1
    for (int k = 0; k < 23; ++k) f[k] = 0.0f;
    f[23] = -1.0f;
    return;
  }
  const int a = G.hdr.y;
  const v3 offa = f3(G.g1);
  const v3 xa = f3(pose[2 * a]);
  const q4 qa = f4q(pose[2 * a + 1]);
  const v3 ra = qrotate(qa, offa);
  const v3 va = add(f3(vel[2 * a]), cross(f3(vel[2 * a + 1]), ra));
  const float ma = bstat[a].x;
  v3 n, offb, vb_n;  // vb_n: the relative velocity in the rows' (n, t, b) basis
  float C0, inv_mb;
  if (kind == kGFloor) {
    // rows: +y normal, world x and z tangents; the floor point is G.g0
    n = mk3(0.0f, 1.0f, 0.0f);
    offb = mk3(0.0f, 0.0f, 0.0f);
    C0 = xa.y + ra.y - G.g0.y;
    vb_n = mk3(va.y, va.x, va.z);
    inv_mb = 0.0f;
  } else {
    const int b = G.hdr.z;
    n = f3(G.g0);
    offb = f3(G.g2);
    const v3 xb = f3(pose[2 * b]);
    const q4 qb = f4q(pose[2 * b + 1]);
    const v3 rb = qrotate(qb, offb);
    const v3 d = sub(add(xa, ra), add(xb, rb));
    const v3 vb = add(f3(vel[2 * b]), cross(f3(vel[2 * b + 1]), rb));
    const v3 vr = sub(va, vb);
    const v3 t = orthonormal_basis(n), bt = cross(n, t);
    C0 = dot(n, d);
    vb_n = mk3(dot(n, vr), dot(t, vr), dot(bt, vr));
    const float mb = bstat[b].x;
    inv_mb = mb > 0.0f ? 1.0f / mb : 0.0f;
  }
  f[0] = n.x; f[1] = n.y; f[2] = n.z;
  f[3] = offa.x; f[4] = offa.y; f[5] = offa.z;
  f[6] = offb.x; f[7] = offb.y; f[8] = offb.z;
  f[9] = C0;
  f[10] = vb_n.x; f[11] = vb_n.y; f[12] = vb_n.z;
  f[13] = ma > 0.0f ? 1.0f / ma : 0.0f;
  f[14] = inv_mb;
  f[15] = slog(S.s0.x); f[16] = slog(S.s0.y); f[17] = slog(S.s0.z);
  f[18] = (float)__float_as_int(S.s2.z);
  f[19] = log10f(fmaxf(S.s1.x, 1.0f)); f[20] = log10f(fmaxf(S.s1.y, 1.0f));
  f[21] = log10f(fmaxf(S.s1.z, 1.0f));
  f[22] = G.g1.w;
  f[23] = kind == kGFloor ? 1.0f : 0.0f;
}

void launch_contact_features(hipStream_t st, int n, int g_first, const GroupGeo* geo,
                             const GroupSt* gst, const float4* pose, const float4* vel,
                             const float4* bstat, float* feat) {
  if (n > 0)
    hipLaunchKernelGGL(k_contact_features, dim3((n + 255) / 256), dim3(256), 0, st, n, g_first,
                       geo, gst, pose, vel, bstat, feat);
}

// The plain warm start of a group, from its features (cached state):
// base[0..2] = slog(ag * lambda), base[3] = log10 of the decayed, clamped k_n.
__device__ __forceinline__ void plain_start(const float* f, float ag, float gamma, float base[4]) {
  for (int r = 0; r < 3; ++r) base[r] = slog(islog(f[15 + r]) * ag);
  const float k = fminf(fmaxf(exp10f(f[19]) * gamma, kPenaltyMin), kPenaltyMax);
  base[3] = log10f(k);
}

// Targets of the logged samples: the groups' dual state at the end of the
// substep, as corrections to the plain warm start.
__global__ void __launch_bounds__(256)
k_contact_targets(int n, int g_first, const GroupGeo* __restrict__ geo,
                  const GroupSt* __restrict__ gst, const float* __restrict__ feat, float ag,
                  float gamma, float* __restrict__ targ, int* __restrict__ ymask) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const GroupSt S = gst[g_first + e];
  float base[4];
  plain_start(feat + (size_t)e * 24, ag, gamma, base);
  float* y = targ + (size_t)e * 4;
  y[0] = slog(S.s0.x) - base[0];
  y[1] = slog(S.s0.y) - base[1];
  y[2] = slog(S.s0.z) - base[2];
  y[3] = log10f(fmaxf(S.s1.x, 1.0f)) - base[3];
  const int h = geo[g_first + e].hdr.x;
  ymask[e] = g_kind(h) == kGPin ? 0 : (g_nrows(h) == 3 ? 15 : 9);
}

void launch_contact_targets(hipStream_t st, int n, int g_first, const GroupGeo* geo,
                            const GroupSt* gst, const float* feat, float ag, float gamma,
                            float* targ, int* ymask) {
  if (n > 0)
    hipLaunchKernelGGL(k_contact_targets, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       g_first, geo, gst, feat, ag, gamma, targ, ymask);
}

// ---------------------------------------------------------------------------
// inference: MFMA bf16
// ---------------------------------------------------------------------------
constexpr int kIn = 24, kHid = 64, kOut = 4;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kHid, kOffW2 = kOffB1 + kHid,
              kOffB2 = kOffW2 + kHid * kHid, kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kOut,
              kNumWeights = kOffB3 + kOut, kOffXm = kNumWeights, kOffXs = kOffXm + kIn,
              kOffYm = kOffXs + kIn, kOffYs = kOffYm + kOut, kNumParams = kOffYs + kOut;

// Largest correction applied (slog units for lambda: a factor e^4 ~ 55;
// decades for k); a larger output is taken as out of the training
// distribution and the plain warm start is kept.
constexpr float kMaxCorrection = 4.0f;

int mlp_num_weights() { return kNumWeights; }
int mlp_num_params() { return kNumParams; }

// 4 wavefronts x 16 contacts per tile of 64; a workgroup stages the weights
// in LDS once (bf16, laid out as the MFMA B operands read them) and walks
// tiles blockIdx.x, blockIdx.x + gridDim.x, ...
//   mode 0: apply the correction to the groups' initial dual state (warm
//           start); a non-finite output leaves the plain warm start
//   mode 2: as mode 0, only on groups whose plain start is empty (normal
//           lambda 0)
//   mode 1: write the 4 network outputs (de-normalised) to `yout`
__global__ void __launch_bounds__(256)
k_mlp_infer(int n, const float* __restrict__ feat, const float* __restrict__ params,
            int g_first, GroupSt* __restrict__ gst, const GroupGeo* __restrict__ geo, float ag,
            float gamma, int mode, float* __restrict__ yout) {
  __shared__ __bf16 w1[32][64];   // rows 24-31 zero (K padding)
  __shared__ __bf16 w2[64][64];
  __shared__ __bf16 w3[64][16];   // columns 4-15 zero
  __shared__ float bias[64 + 64 + 16];
  __shared__ float norm[kIn * 2 + kOut * 2];
  __shared__ __bf16 act[4][16][64 + 8];  // per wavefront: layer input (bf16)
  const int tid = threadIdx.x, wave = tid / 64, l = tid % 64;
  for (int k = tid; k < 32 * 64; k += 256) {
    const int r = k / 64, c = k % 64;
    w1[r][c] = (__bf16)(r < kIn ? params[kOffW1 + r * kHid + c] : 0.0f);
  }
  for (int k = tid; k < 64 * 64; k += 256) w2[k / 64][k % 64] = (__bf16)params[kOffW2 + k];
  for (int k = tid; k < 64 * 16; k += 256) {
    const int r = k / 16, c = k % 16;
    w3[r][c] = (__bf16)(c < kOut ? params[kOffW3 + r * kOut + c] : 0.0f);
  }
  for (int k = tid; k < 144; k += 256)
    bias[k] = k < 64 ? params[kOffB1 + k] : k < 128 ? params[kOffB2 + k - 64]
                                                    : (k - 128 < kOut ? params[kOffB3 + k - 128] : 0.f);
  for (int k = tid; k < kIn * 2 + kOut * 2; k += 256) norm[k] = params[kOffXm + k];
  const int n_tiles = (n + 63) / 64;
  const int row = l % 16, kq = l / 16;  // A row / B column within a tile; K quarter
  // The trip count is the same for every wavefront of the workgroup, so all
  // reach every barrier; rows past the last sample compute on zeros and
  // store nothing.
  for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    __syncthreads();  // weights staged / previous tile's reads of act done
    const int s0 = (tile * 4 + wave) * 16;
    // layer input: normalised features, bf16, K padded to 32
    for (int k = l; k < 16 * 32; k += 64) {
      const int s = k / 32, c = k % 32;
      float v = 0.0f;
      if (c < kIn && s0 + s < n) v = (feat[(size_t)(s0 + s) * kIn + c] - norm[c]) / norm[kIn + c];
      act[wave][s][c] = (__bf16)v;
    }
    __syncthreads();
    floatx4 h[4];
    // ---- layer 1: [16 x 32] x [32 x 64] ----
    {
      bf16x8 a;
      for (int j = 0; j < 8; ++j) a[j] = act[wave][row][8 * kq + j];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        bf16x8 b;
        for (int j = 0; j < 8; ++j) b[j] = w1[8 * kq + j][t * 16 + row];
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
      }
    }
    __syncthreads();
    // SiLU, back to LDS as the next layer's input: D[4*kq + r][t*16 + row]
#pragma unroll
    for (int t = 0; t < 4; ++t)
      for (int r = 0; r < 4; ++r) {
        const int c = t * 16 + row;
        act[wave][4 * kq + r][c] = (__bf16)silu(h[t][r] + bias[c]);
      }
    __syncthreads();
    // ---- layer 2: [16 x 64] x [64 x 64] ----
    floatx4 h2[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f},
                     {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a;
      for (int j = 0; j < 8; ++j) a[j] = act[wave][row][32 * ks + 8 * kq + j];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        bf16x8 b;
        for (int j = 0; j < 8; ++j) b[j] = w2[32 * ks + 8 * kq + j][t * 16 + row];
        h2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, h2[t], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t)
      for (int r = 0; r < 4; ++r) {
        const int c = t * 16 + row;
        act[wave][4 * kq + r][c] = (__bf16)silu(h2[t][r] + bias[64 + c]);
      }
    __syncthreads();
    // ---- layer 3: [16 x 64] x [64 x 16] (4 used) ----
    floatx4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a, b;
      for (int j = 0; j < 8; ++j) {
        a[j] = act[wave][row][32 * ks + 8 * kq + j];
        b[j] = w3[32 * ks + 8 * kq + j][row];
      }
      y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, y, 0, 0, 0);
    }
    // lane holds outputs (sample 4*kq + r, column row)
    if (row < kOut) {
      for (int r = 0; r < 4; ++r) {
        const int s = s0 + 4 * kq + r;
        if (s >= n) continue;
        const float v = (y[r] + bias[128 + row]) * norm[2 * kIn + kOut + row] + norm[2 * kIn + row];
        if (mode == 1) {
          yout[(size_t)s * kOut + row] = v;
          continue;
        }
        // zero, non-finite or beyond the trusted range: keep the plain warm
        // start as it is
        if (v == 0.0f || !(fabsf(v) <= kMaxCorrection)) continue;
        const float* f = feat + (size_t)s * kIn;
        GroupSt* S = gst + g_first + s;
        const int h = geo[g_first + s].hdr.x;
        if (g_kind(h) == kGPin || (mode == 2 && f[15] != 0.0f)) continue;
        const int nr = g_nrows(h);
        if (row == 3) {
          // start at 10^(base + v) after the prelude multiplies k by gamma
          const float kp = fminf(fmaxf(exp10f(f[19]) * gamma, kPenaltyMin), kPenaltyMax);
          const float lk = fminf(fmaxf(log10f(kp) + v, 6.0f), 9.0f);
          S->s1.x = exp10f(lk) / gamma;
        } else if (row == 0 || nr == 3) {
          // start at islog(base + v) after the prelude multiplies lambda by ag
          float lam = islog(slog(islog(f[15 + row]) * ag) + v) / ag;
          if (row == 0) lam = fminf(lam, 0.0f);  // the normal force is compressive (<= 0)
          if (row == 0) S->s0.x = lam;
          else if (row == 1) S->s0.y = lam;
          else S->s0.z = lam;
        }
      }
    }
  }
}

// tiles per workgroup: enough workgroups to fill 256 CUs, each staging the
// 14 KB of weights once for several tiles
void launch_mlp_infer(hipStream_t st, int n, const float* feat, const float* params, int g_first,
                      GroupSt* gst, const GroupGeo* geo, float ag, float gamma, int mode,
                      float* yout) {
  if (n <= 0) return;
  const int tiles = (n + 63) / 64;
  hipLaunchKernelGGL(k_mlp_infer, dim3(std::min(tiles, 1024)), dim3(256), 0, st, n, feat, params,
                     g_first, gst, geo, ag, gamma, mode, yout);
}

// ---------------------------------------------------------------------------
// training: forward and backward on MFMA (bf16 operands, fp32 accumulation)
// ---------------------------------------------------------------------------
// One workgroup (4 wavefronts) per tile of 64 samples; sample s of the batch
// is row idx[s] of X, Y, ymask (row s when idx is null). Loss = sum over
// samples and used outputs of h(yhat - y) * inv_total, y normalised, with
// h(d) = d^2 (huber <= 0) or the Huber loss d^2 for |d| <= huber and
// huber (2|d| - huber) beyond; mask bit o of ymask says whether output o is
// used.
//
// Every product runs as v_mfma_f32_16x16x32_bf16 except the 4-wide layer-3
// backward (VALU). Rounded to bf16 as MFMA operands: the weights, the
// normalised input, both hidden activations (as in inference), the output
// error dy and the pre-activation gradients dz2, dz1. The silu'
// factors, the bias gradients (sums of the unrounded dy, dz2, dz1) and all
// accumulation stay fp32.
//
// MFMA 16x16x32 operand layout, lane l = 16 kq + c (kq = l / 16, c = l % 16):
// A [16 x 32]: a[j] = A[c][8 kq + j]; B [32 x 16]: b[j] = B[8 kq + j][c];
// D [16 x 16]: d[r] = D[4 kq + r][c].
__device__ __forceinline__ floatx4 mfma16(const bf16x8& a, const bf16x8& b, floatx4 acc) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}

// Each workgroup writes its gradient and loss, with plain stores, to its own
// row of `part` (kPartStride floats: the weight gradient, then the loss);
// k_mlp_part_reduce then sums the rows in workgroup order, so the result does
// not depend on scheduling.
constexpr int kPartStride = ((kNumWeights + 1 + 63) / 64) * 64;
int mlp_part_stride() { return kPartStride; }

__global__ void __launch_bounds__(256)
k_mlp_train_tile(int n, const int* __restrict__ idx, const float* __restrict__ X,
                 const float* __restrict__ Y, const int* __restrict__ ymask,
                 const float* __restrict__ params, float inv_total, float huber,
                 float* __restrict__ part) {
  __shared__ __bf16 w1[32][64];    // W1 [in][hid], rows 24-31 zero
  __shared__ __bf16 w2[64][64];    // W2 [hid][hid]
  __shared__ __bf16 w3[64][16];    // W3 [hid][out], columns 4-15 zero
  __shared__ float bias[64 + 64 + 4];
  __shared__ __bf16 xs[64][40];    // layer-1 input [sample][in], columns 24-39 zero
  __shared__ __bf16 a1s[64][72];   // layer-2 input [sample][hid]
  __shared__ __bf16 a2s[64][72];   // layer-3 input [sample][hid]
  __shared__ float dyf[64][4];     // dy [sample][out], fp32
  __shared__ __bf16 dyb[64][16];   // dy as an MFMA operand, columns 4-15 zero
  __shared__ __bf16 dz2b[64][72];  // dz2 [sample][hid]
  __shared__ __bf16 dz1b[64][72];  // dz1 [sample][hid]
  __shared__ int rows[64];
  __shared__ float bpart[2][16][64];  // bias-gradient partials of b2, b1 per (wavefront, kq)
  __shared__ float lpart[16][4];      // loss partials per (wavefront, kq) and output
  float* __restrict__ grad = part + (size_t)blockIdx.x * kPartStride;
  const int tid = threadIdx.x, wave = tid / 64, l = tid % 64;
  const int c = l % 16, kq = l / 16;
  for (int k = tid; k < 32 * 64; k += 256) {
    const int r = k / 64, col = k % 64;
    w1[r][col] = (__bf16)(r < kIn ? params[kOffW1 + r * kHid + col] : 0.0f);
  }
  for (int k = tid; k < 64 * 64; k += 256) w2[k / 64][k % 64] = (__bf16)params[kOffW2 + k];
  for (int k = tid; k < 64 * 16; k += 256) {
    const int r = k / 16, col = k % 16;
    w3[r][col] = (__bf16)(col < kOut ? params[kOffW3 + r * kOut + col] : 0.0f);
  }
  for (int k = tid; k < 132; k += 256)
    bias[k] = k < 64 ? params[kOffB1 + k] : k < 128 ? params[kOffB2 + k - 64] : params[kOffB3 + k - 128];
  const int t0 = blockIdx.x * 64;
  if (tid < 64) rows[tid] = t0 + tid < n ? (idx ? idx[t0 + tid] : t0 + tid) : -1;
  for (int k = tid; k < 64 * 16; k += 256) dyb[k / 16][k % 16] = (__bf16)0.0f;
  __syncthreads();
  for (int k = tid; k < 64 * 40; k += 256) {
    const int s = k / 40, col = k % 40;
    float v = 0.0f;
    if (col < kIn && rows[s] >= 0)
      v = (X[(size_t)rows[s] * kIn + col] - params[kOffXm + col]) / params[kOffXs + col];
    xs[s][col] = (__bf16)v;
  }
  __syncthreads();
  const int s0 = wave * 16;  // this wavefront's samples in the forward pass
  // ---- forward, layer 1: [16 x 32] x [32 x 64] ----
  floatx4 sg1[4], sg2[4];  // silu'(z) of this lane's outputs D[4 kq + r][16 t + c]
  {
    bf16x8 a;
    for (int j = 0; j < 8; ++j) a[j] = xs[s0 + c][8 * kq + j];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      bf16x8 b;
      for (int j = 0; j < 8; ++j) b[j] = w1[8 * kq + j][16 * t + c];
      floatx4 z = mfma16(a, b, floatx4{0.f, 0.f, 0.f, 0.f});
      for (int r = 0; r < 4; ++r) {
        const float zz = z[r] + bias[16 * t + c];
        a1s[s0 + 4 * kq + r][16 * t + c] = (__bf16)silu(zz);
        sg1[t][r] = silu_grad(zz);
      }
    }
  }
  __syncthreads();
  // ---- layer 2: [16 x 64] x [64 x 64] ----
  {
    floatx4 z[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f},
                    {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a;
      for (int j = 0; j < 8; ++j) a[j] = a1s[s0 + c][32 * ks + 8 * kq + j];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        bf16x8 b;
        for (int j = 0; j < 8; ++j) b[j] = w2[32 * ks + 8 * kq + j][16 * t + c];
        z[t] = mfma16(a, b, z[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
      for (int r = 0; r < 4; ++r) {
        const float zz = z[t][r] + bias[64 + 16 * t + c];
        a2s[s0 + 4 * kq + r][16 * t + c] = (__bf16)silu(zz);
        sg2[t][r] = silu_grad(zz);
      }
  }
  __syncthreads();
  // ---- layer 3: [16 x 64] x [64 x 16] (4 used), the loss and dy ----
  {
    floatx4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a, b;
      for (int j = 0; j < 8; ++j) {
        a[j] = a2s[s0 + c][32 * ks + 8 * kq + j];
        b[j] = w3[32 * ks + 8 * kq + j][c];
      }
      y = mfma16(a, b, y);
    }
    if (c < kOut) {
      float lp = 0.0f;
      for (int r = 0; r < 4; ++r) {
        const int s = s0 + 4 * kq + r, row = rows[s];
        float d = 0.0f;
        if (row >= 0 && (ymask[row] >> c & 1)) {
          const float yh = y[r] + bias[128 + c];
          const float yt = (Y[(size_t)row * kOut + c] - params[kOffYm + c]) / params[kOffYs + c];
          d = yh - yt;
          if (huber > 0.0f && fabsf(d) > huber) {
            lp += huber * (2.0f * fabsf(d) - huber) * inv_total;
            d = copysignf(huber, d);
          } else {
            lp += d * d * inv_total;
          }
        }
        const float g = 2.0f * d * inv_total;
        dyf[s][c] = g;
        dyb[s][c] = (__bf16)g;
      }
      lpart[4 * wave + kq][c] = lp;
    }
  }
  __syncthreads();
  // ---- dW3 = a2^T dy: [64 x 64 samples] x [64 samples x 16]; block of wavefront w: rows 16w.. ----
  {
    floatx4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a, b;
      for (int j = 0; j < 8; ++j) {
        a[j] = a2s[32 * ks + 8 * kq + j][16 * wave + c];
        b[j] = dyb[32 * ks + 8 * kq + j][c];
      }
      d = mfma16(a, b, d);
    }
    if (c < kOut)
      for (int r = 0; r < 4; ++r) grad[kOffW3 + (16 * wave + 4 * kq + r) * kOut + c] = d[r];
    if (wave == 0 && l < kOut) {
      float g = 0.0f;
      for (int s = 0; s < 64; ++s) g += dyf[s][l];
      grad[kOffB3 + l] = g;
    }
  }
  // ---- dz2 = (dy W3^T) * silu'(z2) for this wavefront's samples (VALU, K = 4) ----
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int u = 16 * t + c;
    float bsum = 0.0f;
    for (int r = 0; r < 4; ++r) {
      const int s = s0 + 4 * kq + r;
      float h = 0.0f;
      for (int o = 0; o < kOut; ++o) h += dyf[s][o] * (float)w3[u][o];
      const float dz = h * sg2[t][r];
      dz2b[s][u] = (__bf16)dz;
      bsum += dz;
    }
    bpart[0][4 * wave + kq][u] = bsum;
  }
  __syncthreads();
  // ---- dW2 = a1^T dz2: row block 16w, column blocks 0-3 ----
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    floatx4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a, b;
      for (int j = 0; j < 8; ++j) {
        a[j] = a1s[32 * ks + 8 * kq + j][16 * wave + c];
        b[j] = dz2b[32 * ks + 8 * kq + j][16 * t + c];
      }
      d = mfma16(a, b, d);
    }
    for (int r = 0; r < 4; ++r) grad[kOffW2 + (16 * wave + 4 * kq + r) * kHid + 16 * t + c] = d[r];
  }
  // ---- dz1 = (dz2 W2^T) * silu'(z1) for this wavefront's samples ----
  {
    floatx4 h[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f},
                    {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a;
      for (int j = 0; j < 8; ++j) a[j] = dz2b[s0 + c][32 * ks + 8 * kq + j];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        bf16x8 b;  // B[u][i] = W2[i][u]
        for (int j = 0; j < 8; ++j) b[j] = w2[16 * t + c][32 * ks + 8 * kq + j];
        h[t] = mfma16(a, b, h[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = 16 * t + c;
      float bsum = 0.0f;
      for (int r = 0; r < 4; ++r) {
        const float dz = h[t][r] * sg1[t][r];
        dz1b[s0 + 4 * kq + r][i] = (__bf16)dz;
        bsum += dz;
      }
      bpart[1][4 * wave + kq][i] = bsum;
    }
  }
  __syncthreads();
  // ---- dW1 = x^T dz1: column block w, row blocks 0-1 (rows 24-31 dropped) ----
#pragma unroll
  for (int ib = 0; ib < 2; ++ib) {
    floatx4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 a, b;
      for (int j = 0; j < 8; ++j) {
        a[j] = xs[32 * ks + 8 * kq + j][16 * ib + c];
        b[j] = dz1b[32 * ks + 8 * kq + j][16 * wave + c];
      }
      d = mfma16(a, b, d);
    }
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * ib + 4 * kq + r;
      if (i < kIn) grad[kOffW1 + i * kHid + 16 * wave + c] = d[r];
    }
  }
  // bias gradients and the loss: partials summed in a fixed order
  if (tid < 128) {
    const int which = tid / 64, u = tid % 64;
    float g = 0.0f;
    for (int q = 0; q < 16; ++q) g += bpart[which][q][u];
    grad[(which == 0 ? kOffB2 : kOffB1) + u] = g;
  } else if (tid == 128) {
    float lsum = 0.0f;
    for (int q = 0; q < 16; ++q)
      for (int o = 0; o < kOut; ++o) lsum += lpart[q][o];
    grad[kNumWeights] = lsum;
  }
}

// grad[i] = sum of the workgroups' rows (i < kNumWeights); *loss = sum of
// their losses. A workgroup takes 64 entries; its four wavefronts sum the
// rows q, q + 4, q + 8, ... (q = wavefront), then wavefront 0 adds the four
// partial sums in order: a fixed order for every call.
__global__ void __launch_bounds__(256)
k_mlp_part_reduce(int n_blocks, const float* __restrict__ part, float* __restrict__ grad,
                  float* __restrict__ loss) {
  __shared__ float acc[4][64];
  const int q = threadIdx.x / 64, l = threadIdx.x % 64;
  const int i = blockIdx.x * 64 + l;
  float g = 0.0f;
  if (i <= kNumWeights) {
#pragma unroll 8
    for (int b = q; b < n_blocks; b += 4) g += part[(size_t)b * kPartStride + i];
  }
  acc[q][l] = g;
  __syncthreads();
  if (q != 0 || i > kNumWeights) return;
  const float s = ((acc[0][l] + acc[1][l]) + acc[2][l]) + acc[3][l];
  if (i < kNumWeights) grad[i] = s;
  else *loss = s;
}

__global__ void __launch_bounds__(256)
k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
       float* __restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void launch_mlp_train_tile(hipStream_t st, int n, const int* idx, const float* X, const float* Y,
                           const int* ymask, const float* params, float inv_total, float huber,
                           float* part, float* grad, float* loss) {
  if (n <= 0) return;
  const int blocks = (n + 63) / 64;
  hipLaunchKernelGGL(k_mlp_train_tile, dim3(blocks), dim3(256), 0, st, n, idx, X, Y, ymask,
                     params, inv_total, huber, part);
  hipLaunchKernelGGL(k_mlp_part_reduce, dim3((kNumWeights + 1 + 63) / 64), dim3(256), 0, st,
                     blocks, part, grad, loss);
}

void launch_adam(hipStream_t st, int n, float* w, const float* g, float* m, float* v, float lr,
                 float b1, float b2, float c1, float c2, float eps) {
  if (n > 0)
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, n, w, g, m, v, lr, b1, b2,
                       c1, c2, eps);
}

}  // namespace avbd
