// Learned CG warm start: per-node MLP kernels (BF16 MFMA, FP32 accumulation)
// and the model object of the C ABI.
//
// Network: y = W4 s(W3 s(W2 s(W1 f + b1) + b2) + b3) + b4, s = SiLU, 12 inputs,
// three hidden layers of 128, 3 outputs.
//
// Node features (per free node; v^n, v^{n-1}, v^{n-2} are the velocities of
// the last three steps, i.e. the displacement increments of the reference's
// position update x += dt v, fem_position_update.h:25-39, divided by dt):
//   s      = |v^n| + |v^{n-1}| + |v^{n-2}|
//   f0..8  = (v^n, v^{n-1}, v^{n-2}) / s                 (0 when s = 0)
//   f9..11 = log(1 + cK K_ii,dd / (cM m_i)) / 4            (rest stiffness diagonal)
// CG start from the outputs y:
//   v_start = v^n + y0 (v^n - v^{n-1}) + y1 (v^{n-1} - v^{n-2}) + y2 v^n.
// In normalised form v_start = v^n + s u, u = y0 D1 + y1 D2 + y2 F0 with
// D1 = f0..2 - f3..5, D2 = f3..5 - f6..8, F0 = f0..2.  A new model starts at
// y = (2, -1, 0), quadratic extrapolation of the velocity; y = 0 is the plain
// start.  Training target u* = (v^{n+1} - v^n) / s with loss weight s^2.
//
// MFMA layout (v_mfma_f32_32x32x16_bf16, 64 lanes, lane = 32 h + j):
//   A operand: lane holds A[j][8h .. 8h+7]; B operand: lane holds B[8h .. 8h+7][j];
//   C: register r of lane holds C[8 (r/4) + 4h + r%4][j].
// Hidden units are rows and nodes are columns (32 nodes per wave).  The C
// registers of one layer are, after bias and SiLU, the B operand of the next
// layer without any data movement: K-block kb of the next layer takes
// registers 8 (kb%2) .. 8 (kb%2)+7 of output tile kb/2, which hold hidden units
//   u(kb, h, e) = 32 (kb/2) + 16 (kb%2) + 8 (e/4) + 4h + e%4.
// The weight fragments are stored in that K order (positions built on the
// host), so every operand load is one 16-byte load per lane.
//
// Precision: weights are BF16; inputs, activations and back-propagated
// gradients enter the MFMAs as BF16 hi + lo pairs (about 16 significant
// bits), so the network is a smooth function of its inputs: a plain BF16
// rounding of activations would add per-node noise that the stiff operator
// of the CG system amplifies.
#include "mlp.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "otfem/otfem.h"

namespace otfem {

namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x16 __attribute__((__vector_size__(16 * sizeof(float))));

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int unit_of(int t, int r, int h) {
  return 32 * t + 8 * (r >> 2) + 4 * h + (r & 3);
}

// The logistic 1 / (1 + e^-z) with the hardware reciprocal (v_rcp_f32, 1 ulp)
// instead of an IEEE division: a division costs ten instructions, and the
// network evaluates 384 of them per node.
__device__ __forceinline__ float logistic(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

__device__ __forceinline__ float silu(float z) { return z * logistic(z); }

__device__ __forceinline__ float dsilu(float z) {
  const float s = logistic(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int e = 0; e < 8; ++e) z[e] = (__bf16)0.0f;
  return z;
}

__device__ __forceinline__ f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int r = 0; r < 16; ++r) z[r] = 0.0f;
  return z;
}

__device__ __forceinline__ uint16_t bits(__bf16 b) {
  uint16_t u;
  __builtin_memcpy(&u, &b, 2);
  return u;
}

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi)
__device__ __forceinline__ void split(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// acc[t] = sum_kb W[t][kb] (bh[kb] + bl[kb])   (OT output tiles, KB K-blocks)
template <int OT, int KB>
__device__ __forceinline__ void layer2(const bf16x8 *__restrict__ W, const bf16x8 *bh,
                                       const bf16x8 *bl, f32x16 *acc, int lane) {
#pragma unroll
  for (int t = 0; t < OT; ++t) acc[t] = zero16();
#pragma unroll
  for (int kb = 0; kb < KB; ++kb)
#pragma unroll
    for (int t = 0; t < OT; ++t) {
      const bf16x8 w = W[(t * KB + kb) * 64 + lane];
      acc[t] = mfma(w, bh[kb], acc[t]);
      acc[t] = mfma(w, bl[kb], acc[t]);
    }
}

// bias + SiLU of a 128-unit layer into the next layer's B operands
__device__ __forceinline__ void bias_act(const f32x16 *acc, const float *__restrict__ bias, int h,
                                         bf16x8 *bh, bf16x8 *bl) {
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float a = silu(acc[t][r] + bias[unit_of(t, r, h)]);
      __bf16 hi, lo;
      split(a, hi, lo);
      bh[2 * t + (r >> 3)][r & 7] = hi;
      bl[2 * t + (r >> 3)][r & 7] = lo;
    }
}

// Layer-1 operands from this lane's 8 inputs (f[8h .. 8h+7]).
__device__ __forceinline__ void input_operands(const float *x8, bf16x8 &bh, bf16x8 &bl) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 hi, lo;
    split(x8[e], hi, lo);
    bh[e] = hi;
    bl[e] = lo;
  }
}

// Forward pass of one wave (32 nodes); lanes with h == 0 receive y.
__device__ __forceinline__ void mlp_forward(const bf16x8 *__restrict__ wf,
                                            const float *__restrict__ theta, const float *x8,
                                            int lane, float y[3]) {
  const int h = lane >> 5;
  f32x16 acc[4];
  bf16x8 bh[8], bl[8];
  input_operands(x8, bh[0], bl[0]);
  layer2<4, 1>(wf + kFragF1 / 8, bh, bl, acc, lane);
  bias_act(acc, theta + kOffB1, h, bh, bl);
  layer2<4, 8>(wf + kFragF2 / 8, bh, bl, acc, lane);
  bias_act(acc, theta + kOffB2, h, bh, bl);
  layer2<4, 8>(wf + kFragF3 / 8, bh, bl, acc, lane);
  bias_act(acc, theta + kOffB3, h, bh, bl);
  f32x16 o[1];
  layer2<1, 8>(wf + kFragF4 / 8, bh, bl, o, lane);
  y[0] = o[0][0] + theta[kOffB4];
  y[1] = o[0][1] + theta[kOffB4 + 1];
  y[2] = o[0][2] + theta[kOffB4 + 2];
}

// u = y0 D1 + y1 D2 + y2 F0 from a feature row
__device__ __forceinline__ void start_change(const float *f, const float y[3], float u[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d)
    u[d] = y[0] * (f[d] - f[3 + d]) + y[1] * (f[3 + d] - f[6 + d]) + y[2] * f[d];
}

// ---------------------------------------------------------------------------
// Solver-step kernels

// One wave per 32 nodes: features, optional sample record, optional
// prediction of the CG start, and the history shift v^{n-2} <- v^{n-1} <- v^n.
__global__ __launch_bounds__(256) void k_ws_predict(WsArgs A) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, j = lane & 31;
  const int64_t node = ((int64_t)blockIdx.x * 4 + wave) * 32 + j;
  const bool valid = node < A.N;
  float f[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) f[e] = 0.0f;
  double v[3] = {0.0, 0.0, 0.0};
  float4 h1 = make_float4(0.f, 0.f, 0.f, 0.f), h2 = h1;
  float s = 0.0f;
  bool fixed = true;
  if (valid) {
    v[0] = A.v[3 * node];
    v[1] = A.v[3 * node + 1];
    v[2] = A.v[3 * node + 2];
    h1 = A.hist1[node];
    h2 = A.hist2[node];
    fixed = A.fixed[node] != 0;
    const float v0 = (float)v[0], v1 = (float)v[1], v2 = (float)v[2];
    s = sqrtf(v0 * v0 + v1 * v1 + v2 * v2) + sqrtf(h1.x * h1.x + h1.y * h1.y + h1.z * h1.z) +
        sqrtf(h2.x * h2.x + h2.y * h2.y + h2.z * h2.z);
    if (s > 0.0f && isfinite(s)) {
      const float inv = 1.0f / s;
      f[0] = v0 * inv; f[1] = v1 * inv; f[2] = v2 * inv;
      f[3] = h1.x * inv; f[4] = h1.y * inv; f[5] = h1.z * inv;
      f[6] = h2.x * inv; f[7] = h2.y * inv; f[8] = h2.z * inv;
    } else {
      s = 0.0f;
    }
    if (!fixed) {
      const float4 k = A.kd[node];
      const double q = A.cK / (A.cM * A.mass[node]);
      f[9] = 0.25f * log1pf((float)(q * k.x));
      f[10] = 0.25f * log1pf((float)(q * k.y));
      f[11] = 0.25f * log1pf((float)(q * k.z));
    }
  }
  float x8[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) x8[e] = h ? f[8 + e] : f[e];
  if (A.record && valid && !fixed) {
    const int64_t i = A.rec_base + A.free_rank[node];
    float4 *dst = reinterpret_cast<float4 *>(A.rec_X + 16 * i + 8 * h);
    dst[0] = make_float4(x8[0], x8[1], x8[2], x8[3]);
    dst[1] = make_float4(x8[4], x8[5], x8[6], x8[7]);
    if (h == 0) A.rec_S[i] = s;
  }
  if (A.predict) {  // uniform over the grid: every lane reaches the MFMAs
    float y[3];
    mlp_forward(reinterpret_cast<const bf16x8 *>(A.wf), A.theta, x8, lane, y);
    if (h == 0 && valid) {
      const double hv1[3] = {h1.x, h1.y, h1.z}, hv2[3] = {h2.x, h2.y, h2.z};
      double w[3];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        w[d] = v[d] + (double)y[0] * (v[d] - hv1[d]) + (double)y[1] * (hv1[d] - hv2[d]) +
               (double)y[2] * v[d];
      const bool ok = isfinite(w[0]) && isfinite(w[1]) && isfinite(w[2]);
      if (!ok && !fixed) atomicAdd(A.bad, 1);
#pragma unroll
      for (int d = 0; d < 3; ++d) A.vw[3 * node + d] = (ok && !fixed) ? w[d] : v[d];
    }
  }
  if (h == 0 && valid) {
    A.hist2[node] = h1;
    A.hist1[node] = make_float4((float)v[0], (float)v[1], (float)v[2], 0.0f);
  }
}

// After the step: target u* = (v^{n+1} - v^n) / s and weight s^2.
__global__ __launch_bounds__(256) void k_ws_target(WsArgs A) {
  const int64_t node = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (node >= A.N || A.fixed[node]) return;
  const int64_t i = A.rec_base + A.free_rank[node];
  const float s = A.rec_S[i];
  const float4 vn = A.hist1[node];
  float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
  if (s > 0.0f) {
    const double inv = 1.0 / (double)s;
    t.x = (float)((A.v[3 * node] - (double)vn.x) * inv);
    t.y = (float)((A.v[3 * node + 1] - (double)vn.y) * inv);
    t.z = (float)((A.v[3 * node + 2] - (double)vn.z) * inv);
    t.w = s * s;
  }
  A.rec_T[i] = t;
}

// ---------------------------------------------------------------------------
// Stand-alone forward over feature rows (FP32 [n][16]) -> FP32 [n][3].
__global__ __launch_bounds__(256) void k_mlp_rows(const bf16x8 *wf, const float *theta,
                                                  const float *X, int64_t n, float *Y) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, j = lane & 31;
  const int64_t c = ((int64_t)blockIdx.x * 4 + wave) * 32 + j;
  float x8[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) x8[e] = c < n ? X[16 * c + 8 * h + e] : 0.0f;
  float y[3];
  mlp_forward(wf, theta, x8, lane, y);
  if (h == 0 && c < n) {
    Y[3 * c] = y[0];
    Y[3 * c + 1] = y[1];
    Y[3 * c + 2] = y[2];
  }
}

// ---------------------------------------------------------------------------
// Training: forward per sample column, loss gradient dL/dy, backward per
// column, weight gradients as a reduction over columns, Adam.
//
// Per-column arrays are blocked by 32 columns: element (u, c) of an array with
// U rows sits at ((c / 32) U + u) 32 + c % 32 (blk() below), so a wave's 32
// columns of one row are contiguous and row offsets are constants.

__device__ __forceinline__ int64_t blk(int U, int u, int64_t c) {
  return ((c >> 5) * U + u) * 32 + (c & 31);
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

enum { PICK_RANDOM = 0, PICK_CONTIGUOUS = 1, PICK_STEPS = 2 };

struct FwdArgs {
  const bf16x8 *wf;
  const float *theta;
  const float *X;      // [count][16]
  const float4 *T;     // [count] target, weight
  int64_t count;
  int pick;            // PICK_*
  uint64_t seed;
  int64_t step;
  int64_t base;        // PICK_CONTIGUOUS: first sample
  const int32_t *steps;// PICK_STEPS: recorded step of each group of n_free columns
  int64_t n_free;
  int64_t batch;       // columns (multiple of 128)
  int64_t nvalid;      // columns < nvalid carry a sample
  int mse;             // write dL/dy of the weighted MSE loss (else only y)
  float inv_w;         // 1 / mean loss weight (MSE)
  float *Xt;           // U = 16 inputs
  float *Z;            // 3 layers x (U = 128) pre-activations
  float *H;            // 3 layers x (U = 128) activations
  float *Y;            // U = 4 outputs
  float *G4;           // U = 4: dL/dy
  double *loss_part;   // [gridDim.x]
};

__device__ __forceinline__ int64_t sample_of(const FwdArgs &A, int64_t c) {
  if (A.pick == PICK_CONTIGUOUS) return A.base + c;
  if (A.pick == PICK_STEPS) return (int64_t)A.steps[c / A.n_free] * A.n_free + c % A.n_free;
  return (int64_t)(mix64(A.seed ^ mix64((uint64_t)A.step * 0x100000001b3ull + c)) %
                   (uint64_t)A.count);
}

// Z, H point at this lane's (unit 4h, column c) element of a layer.
__device__ __forceinline__ void train_act(const f32x16 *acc, const float *__restrict__ bias, int h,
                                          bf16x8 *bh, bf16x8 *bl, float *Z, float *H) {
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int du = unit_of(t, r, 0);
      const float z = acc[t][r] + bias[du + 4 * h];
      const float a = silu(z);
      Z[du * 32] = z;
      H[du * 32] = a;
      __bf16 hi, lo;
      split(a, hi, lo);
      bh[2 * t + (r >> 3)][r & 7] = hi;
      bl[2 * t + (r >> 3)][r & 7] = lo;
    }
}

__global__ __launch_bounds__(256) void k_mlp_fwd(FwdArgs A) {
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, j = lane & 31;
  const int64_t B = A.batch;
  const int64_t c = ((int64_t)blockIdx.x * 4 + wave) * 32 + j;
  const bool valid = c < A.nvalid;
  const int64_t idx = valid ? sample_of(A, c) : 0;
  float x8[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) x8[e] = valid ? A.X[16 * idx + 8 * h + e] : 0.0f;
#pragma unroll
  for (int e = 0; e < 8; ++e) A.Xt[blk(16, 8 * h + e, c)] = x8[e];

  const int64_t off = blk(128, 4 * h, c);
  f32x16 acc[4];
  bf16x8 bh[8], bl[8];
  input_operands(x8, bh[0], bl[0]);
  layer2<4, 1>(A.wf + kFragF1 / 8, bh, bl, acc, lane);
  train_act(acc, A.theta + kOffB1, h, bh, bl, A.Z + off, A.H + off);
  layer2<4, 8>(A.wf + kFragF2 / 8, bh, bl, acc, lane);
  train_act(acc, A.theta + kOffB2, h, bh, bl, A.Z + 128 * B + off, A.H + 128 * B + off);
  layer2<4, 8>(A.wf + kFragF3 / 8, bh, bl, acc, lane);
  train_act(acc, A.theta + kOffB3, h, bh, bl, A.Z + 256 * B + off, A.H + 256 * B + off);
  f32x16 o[1];
  layer2<1, 8>(A.wf + kFragF4 / 8, bh, bl, o, lane);

  double lsum = 0.0;
  if (h == 0) {
    float y[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      y[r] = o[0][r] + A.theta[kOffB4 + r];
      A.Y[blk(4, r, c)] = y[r];
    }
    if (A.mse) {
      float g[3] = {0.f, 0.f, 0.f};
      if (valid) {
        float f[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) f[e] = A.X[16 * idx + e];
        const float4 t = A.T[idx];
        const float w = t.w * A.inv_w, invn = 1.0f / (float)A.nvalid;
        float u[3];
        start_change(f, y, u);
        const float d[3] = {u[0] - t.x, u[1] - t.y, u[2] - t.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) lsum += 0.5 * (double)w * (double)d[k] * (double)d[k];
        lsum /= (double)A.nvalid;
        // dL/dy_k = w/n (d . B_k), B = (D1, D2, F0)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          g[0] += d[k] * (f[k] - f[3 + k]);
          g[1] += d[k] * (f[3 + k] - f[6 + k]);
          g[2] += d[k] * f[k];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) g[r] *= w * invn;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) A.G4[blk(4, r, c)] = g[r];
    }
  }
#pragma unroll
  for (int o2 = 32; o2 > 0; o2 >>= 1) lsum += __shfl_xor(lsum, o2, 64);
  if (lane == 0) red[wave] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) A.loss_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

struct BwdArgs {
  const bf16x8 *wb;
  int64_t batch;
  const float *Z;
  const float *G4;
  float *G;  // 3 layers x (U = 128): dL/dZ of the hidden layers
};

__device__ __forceinline__ void train_grad(const f32x16 *acc, const float *Z, float *G,
                                           bf16x8 *bh, bf16x8 *bl) {
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int du = unit_of(t, r, 0);
      const float g = acc[t][r] * dsilu(Z[du * 32]);
      G[du * 32] = g;
      __bf16 hi, lo;
      split(g, hi, lo);
      bh[2 * t + (r >> 3)][r & 7] = hi;
      bl[2 * t + (r >> 3)][r & 7] = lo;
    }
}

__global__ __launch_bounds__(256) void k_mlp_bwd(BwdArgs A) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, j = lane & 31;
  const int64_t B = A.batch;
  const int64_t c = ((int64_t)blockIdx.x * 4 + wave) * 32 + j;
  // dL/dy as a B operand: units 0..2 sit at (kb 0, h 0, e 0..2)
  bf16x8 gh = zero8(), gl = zero8();
  if (h == 0) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      __bf16 hi, lo;
      split(A.G4[blk(4, r, c)], hi, lo);
      gh[r] = hi;
      gl[r] = lo;
    }
  }
  const int64_t off = blk(128, 4 * h, c);
  f32x16 acc[4];
  bf16x8 bh[8], bl[8];
  layer2<4, 1>(A.wb + kFragB4 / 8, &gh, &gl, acc, lane);             // W4^T dL/dy
  train_grad(acc, A.Z + 256 * B + off, A.G + 256 * B + off, bh, bl);  // G3
  layer2<4, 8>(A.wb + kFragB3 / 8, bh, bl, acc, lane);               // W3^T G3
  train_grad(acc, A.Z + 128 * B + off, A.G + 128 * B + off, bh, bl);  // G2
  layer2<4, 8>(A.wb + kFragB2 / 8, bh, bl, acc, lane);               // W2^T G2
  train_grad(acc, A.Z + off, A.G + off, bh, bl);                      // G1
}

struct GradArgs {
  const float *G;
  const float *G4;
  const float *H;
  const float *Xt;
  int64_t batch;
  int64_t chunk;
  float *part;  // [gridDim.x][kMlpParams]
};

// 8 consecutive FP32 values as hi + lo BF16 operands (and their sum)
__device__ __forceinline__ void load_split(const float *p, bf16x8 &hi, bf16x8 &lo, float *sum) {
  const float4 a = reinterpret_cast<const float4 *>(p)[0];
  const float4 b = reinterpret_cast<const float4 *>(p)[1];
  const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 x, y;
    split(v[e], x, y);
    hi[e] = x;
    lo[e] = y;
    if (sum) *sum += v[e];
  }
}

// acc += (ah + al)(bh + bl) without the lo*lo term
__device__ __forceinline__ f32x16 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f32x16 acc) {
  acc = mfma(ah, bh, acc);
  acc = mfma(al, bh, acc);
  return mfma(ah, bl, acc);
}

// dW = G H^T and db = sum G over one chunk of columns, per workgroup.
// Wave w owns output rows 32w .. 32w+31 of W1, W2, W3 and columns 32w ..
// 32w+31 of W4.
__global__ __launch_bounds__(256) void k_mlp_wgrad(GradArgs A) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, i = lane & 31;
  const int64_t B = A.batch;
  const int64_t c0 = (int64_t)blockIdx.x * A.chunk;
  const int64_t c1 = c0 + A.chunk < B ? c0 + A.chunk : B;
  float *P = A.part + (int64_t)blockIdx.x * kMlpParams;
  const int row = 32 * wave + i;

  // W2 (G2, H1) and W3 (G3, H2)
  for (int L = 2; L <= 3; ++L) {
    const float *Gl = A.G + (int64_t)(L - 1) * 128 * B;
    const float *Hp = A.H + (int64_t)(L - 2) * 128 * B;
    f32x16 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = zero16();
    float bsum = 0.0f;
    for (int64_t kk = c0; kk < c1; kk += 16) {
      const int64_t k = kk + 8 * h;
      bf16x8 ah, al;
      load_split(Gl + blk(128, row, k), ah, al, &bsum);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        bf16x8 hh, hl;
        load_split(Hp + blk(128, 32 * t + i, k), hh, hl, nullptr);
        acc[t] = mfma3(ah, al, hh, hl, acc[t]);
      }
    }
    const int offW = L == 2 ? kOffW2 : kOffW3, offB = L == 2 ? kOffB2 : kOffB3;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) P[offW + unit_of(wave, r, h) * 128 + 32 * t + i] = acc[t][r];
    bsum += __shfl_xor(bsum, 32, 64);
    if (h == 0) P[offB + row] = bsum;
  }
  // W1 (G1, inputs)
  {
    f32x16 acc = zero16();
    float bsum = 0.0f;
    for (int64_t kk = c0; kk < c1; kk += 16) {
      const int64_t k = kk + 8 * h;
      bf16x8 ah, al;
      load_split(A.G + blk(128, row, k), ah, al, &bsum);
      bf16x8 xh = zero8(), xl = zero8();
      if (i < 16) load_split(A.Xt + blk(16, i, k), xh, xl, nullptr);
      acc = mfma3(ah, al, xh, xl, acc);
    }
    if (i < kMlpIn) {
#pragma unroll
      for (int r = 0; r < 16; ++r) P[kOffW1 + unit_of(wave, r, h) * kMlpIn + i] = acc[r];
    }
    bsum += __shfl_xor(bsum, 32, 64);
    if (h == 0) P[kOffB1 + row] = bsum;
  }
  // W4 (dL/dy rows 0..2, H3)
  {
    const float *H3 = A.H + 256 * B;
    f32x16 acc = zero16();
    float bsum = 0.0f;
    for (int64_t kk = c0; kk < c1; kk += 16) {
      const int64_t k = kk + 8 * h;
      bf16x8 ah = zero8(), al = zero8();
      if (i < 3) load_split(A.G4 + blk(4, i, k), ah, al, &bsum);
      bf16x8 hh, hl;
      load_split(H3 + blk(128, 32 * wave + i, k), hh, hl, nullptr);
      acc = mfma3(ah, al, hh, hl, acc);
    }
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 3; ++r) P[kOffW4 + r * 128 + 32 * wave + i] = acc[r];
    }
    bsum += __shfl_xor(bsum, 32, 64);
    if (wave == 0 && h == 0 && i < 3) P[kOffB4 + i] = bsum;
  }
}

// grad[p] = sum of the workgroup partials, in workgroup order.
__global__ void k_mlp_reduce(const float *part, int nparts, float *grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  double s = 0.0;
  for (int b = 0; b < nparts; ++b) s += part[(int64_t)b * kMlpParams + p];
  grad[p] = (float)s;
}

struct AdamArgs {
  float *theta, *m, *v;
  const float *part;
  int nparts;
  float lr, beta1, beta2, eps, corr1, corr2;
  const int32_t *fpos, *bpos;
  uint16_t *wf, *wb;
  const double *loss_part;
  int nloss;
  float *loss_out;
};

__global__ void k_mlp_adam(AdamArgs A) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && A.loss_out) {
    double s = 0.0;
    for (int b = 0; b < A.nloss; ++b) s += A.loss_part[b];
    *A.loss_out = (float)s;
  }
  if (p >= kMlpParams) return;
  double g = 0.0;
  for (int b = 0; b < A.nparts; ++b) g += A.part[(int64_t)b * kMlpParams + p];
  const float gf = (float)g;
  const float m = A.beta1 * A.m[p] + (1.0f - A.beta1) * gf;
  const float v = A.beta2 * A.v[p] + (1.0f - A.beta2) * gf * gf;
  A.m[p] = m;
  A.v[p] = v;
  const float th = A.theta[p] - A.lr * (m / A.corr1) / (sqrtf(v / A.corr2) + A.eps);
  A.theta[p] = th;
  const uint16_t bb = bits((__bf16)th);
  if (A.fpos[p] >= 0) A.wf[A.fpos[p]] = bb;
  if (A.bpos[p] >= 0) A.wb[A.bpos[p]] = bb;
}

// Sum of the loss weights (column 3 of the targets), per workgroup.
__global__ __launch_bounds__(256) void k_weight_sum(const float4 *T, int64_t n, double *part) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += T[i].w;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ---------------------------------------------------------------------------
// Residual loss.  For a recorded step with free-node samples i, the start
// error in velocity units is e_i = s_i (u_i - u*_i); the plain start's is
// e0_i = -s_i u*_i.  Loss over a minibatch of K steps:
//   L = sum_k |A e_k|^2 / sum_k |A e0_k|^2.

__global__ void k_pick_steps(int32_t *steps, int K, int64_t n_steps, uint64_t seed, int64_t t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K)
    steps[k] = (int32_t)(mix64(seed ^ mix64(((uint64_t)t << 20) + (uint64_t)k + 0x51ed27u)) %
                         (uint64_t)n_steps);
}

// E[c] for the batch columns (or, with Y == nullptr, e0 of every sample)
__global__ __launch_bounds__(256) void k_res_err(const float *Y, const float *X, const float4 *T,
                                                 const float *S, const int32_t *steps,
                                                 int64_t n_free, int64_t n, float4 *E) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int64_t idx = steps ? (int64_t)steps[c / n_free] * n_free + c % n_free : c;
  const float4 t = T[idx];
  const float s = S[idx];
  float u[3] = {0.f, 0.f, 0.f};
  if (Y) {
    float f[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) f[e] = X[16 * idx + e];
    const float y[3] = {Y[blk(4, 0, c)], Y[blk(4, 1, c)], Y[blk(4, 2, c)]};
    start_change(f, y, u);
  }
  E[c] = make_float4(s * (u[0] - t.x), s * (u[1] - t.y), s * (u[2] - t.z), 0.0f);
}

// y[k n + i] = sum_j A_ij x[k n + j] for `nvec` vectors of n block rows.
__global__ __launch_bounds__(256) void k_bsr_spmv(const int64_t *rp, const int32_t *cols,
                                                  const float *blocks, int64_t n, int64_t nvec,
                                                  const float4 *x, float4 *y) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n * nvec) return;
  const int64_t k = g / n, i = g % n;
  const float4 *xv = x + k * n;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
    const float *b = blocks + 9 * q;
    const float4 v = xv[cols[q]];
    a0 += b[0] * v.x + b[1] * v.y + b[2] * v.z;
    a1 += b[3] * v.x + b[4] * v.y + b[5] * v.z;
    a2 += b[6] * v.x + b[7] * v.y + b[8] * v.z;
  }
  y[g] = make_float4(a0, a1, a2, 0.0f);
}

// Per-step |Q|^2 (one workgroup per vector, fixed order).
__global__ __launch_bounds__(256) void k_vec_norms(const float4 *Q, int64_t n, double *out) {
  __shared__ double red[4];
  const float4 *q = Q + (int64_t)blockIdx.x * n;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256)
    s += (double)q[i].x * q[i].x + (double)q[i].y * q[i].y + (double)q[i].z * q[i].z;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// dL/dy of the residual loss for every batch column, and loss partials.
__global__ __launch_bounds__(256) void k_res_grad(const float4 *Q, const float4 *Gv,
                                                  const double *Q0, const int32_t *steps,
                                                  const float *X, const float *S, int64_t n_free,
                                                  int K, int64_t nvalid, int64_t batch, float *G4,
                                                  double *loss_part) {
  __shared__ double red[4];
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  float g[3] = {0.f, 0.f, 0.f};
  double q0 = 0.0;  // the plain start's sum over the minibatch
  for (int k = 0; k < K; ++k) q0 += Q0[steps[k]];
  if (c < nvalid) {
    const int32_t st = steps[c / n_free];
    const int64_t idx = (int64_t)st * n_free + c % n_free;
    if (q0 > 0.0) {
      const float4 q = Q[c], gv = Gv[c];
      l = ((double)q.x * q.x + (double)q.y * q.y + (double)q.z * q.z) / q0;
      // dL/de = 2 A q / q0; dL/du = s dL/de; dL/dy_k = dL/du . B_k
      const float sc = (float)(2.0 / q0) * S[idx];
      const float du[3] = {sc * gv.x, sc * gv.y, sc * gv.z};
      float f[9];
#pragma unroll
      for (int e = 0; e < 9; ++e) f[e] = X[16 * idx + e];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        g[0] += du[d] * (f[d] - f[3 + d]);
        g[1] += du[d] * (f[3 + d] - f[6 + d]);
        g[2] += du[d] * f[d];
      }
    }
  }
  if (c < batch) {
#pragma unroll
    for (int r = 0; r < 3; ++r) G4[blk(4, r, c)] = g[r];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

hipError_t enqueue_ws_predict(const WsArgs &a, hipStream_t s) {
  const int64_t blocks = (a.N + 127) / 128;
  hipLaunchKernelGGL(k_ws_predict, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t enqueue_ws_target(const WsArgs &a, hipStream_t s) {
  const int64_t blocks = (a.N + 255) / 256;
  hipLaunchKernelGGL(k_ws_target, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace otfem

// ---------------------------------------------------------------------------
// Model object (host side)

using namespace otfem;

struct otfem_model {
  std::string err;
  int device = -1;
  hipStream_t stream = nullptr;
  float *theta = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  uint16_t *wf = nullptr, *wb = nullptr;
  int32_t *fpos = nullptr, *bpos = nullptr;
  std::vector<int32_t> fpos_host, bpos_host;
  int64_t adam_t = 0;
  // samples
  int64_t cap = 0, count = 0;
  float *X = nullptr;
  float *S = nullptr;
  float4 *T = nullptr;
  // training scratch for `scratch_batch` columns
  int64_t scratch_batch = 0;
  float *Xt = nullptr, *H = nullptr, *Z = nullptr, *G = nullptr, *G4 = nullptr, *Y = nullptr;
  float *part = nullptr, *grad = nullptr;
  double *loss_part = nullptr;
  float *loss_hist = nullptr;
  int64_t loss_hist_cap = 0;
  double *wsum_part = nullptr;
};

namespace {

constexpr int kWsumBlocks = 256;

otfem_status mfail(otfem_model *m, otfem_status st, const std::string &msg) {
  if (m) m->err = msg;
  return st;
}

#define MCHK(m, call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return mfail((m), OTFEM_HIP_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t malloc_n(T **p, size_t n) {
  return hipMalloc(reinterpret_cast<void **>(p), (n ? n : 1) * sizeof(T));
}

template <typename T>
void free_p(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// device buffer freed at scope exit
template <typename T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() { free_p(p); }
};

uint16_t host_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

// K-slot (kb, h, e) of hidden unit `in` in the order the C registers of the
// producing layer deliver it.
void kslot(int in, int &kb, int &h, int &e) {
  const int t = in / 32, rem = in % 32;
  kb = 2 * t + rem / 16;
  h = (rem % 8) / 4;
  e = 4 * ((rem % 16) / 8) + rem % 4;
}

void build_positions(std::vector<int32_t> &fp, std::vector<int32_t> &bp) {
  fp.assign(kMlpParams, -1);
  bp.assign(kMlpParams, -1);
  int kb, h, e;
  for (int o = 0; o < 128; ++o)
    for (int i = 0; i < kMlpIn; ++i)
      fp[kOffW1 + o * kMlpIn + i] = kFragF1 + ((o / 32) * 64 + (i / 8) * 32 + o % 32) * 8 + i % 8;
  for (int L = 2; L <= 3; ++L) {
    const int off = L == 2 ? kOffW2 : kOffW3, F = L == 2 ? kFragF2 : kFragF3,
              Bk = L == 2 ? kFragB2 : kFragB3;
    for (int o = 0; o < 128; ++o)
      for (int i = 0; i < 128; ++i) {
        kslot(i, kb, h, e);
        fp[off + o * 128 + i] = F + (((o / 32) * 8 + kb) * 64 + h * 32 + o % 32) * 8 + e;
        kslot(o, kb, h, e);
        bp[off + o * 128 + i] = Bk + (((i / 32) * 8 + kb) * 64 + h * 32 + i % 32) * 8 + e;
      }
  }
  for (int o = 0; o < kMlpOut; ++o)
    for (int i = 0; i < 128; ++i) {
      kslot(i, kb, h, e);
      fp[kOffW4 + o * 128 + i] = kFragF4 + (kb * 64 + h * 32 + o) * 8 + e;
      bp[kOffW4 + o * 128 + i] = kFragB4 + ((i / 32) * 64 + i % 32) * 8 + o;
    }
}

otfem_status upload_params(otfem_model *m, const std::vector<float> &theta) {
  std::vector<uint16_t> wf(kFragFwd, 0), wb(kFragBwd, 0);
  for (int p = 0; p < kMlpParams; ++p) {
    if (m->fpos_host[p] >= 0) wf[m->fpos_host[p]] = host_bf16(theta[p]);
    if (m->bpos_host[p] >= 0) wb[m->bpos_host[p]] = host_bf16(theta[p]);
  }
  MCHK(m, hipMemcpyAsync(m->theta, theta.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice,
                         m->stream));
  MCHK(m, hipMemcpyAsync(m->wf, wf.data(), kFragFwd * 2, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(m->wb, wb.data(), kFragBwd * 2, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemsetAsync(m->adam_m, 0, kMlpParams * sizeof(float), m->stream));
  MCHK(m, hipMemsetAsync(m->adam_v, 0, kMlpParams * sizeof(float), m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  m->adam_t = 0;
  return OTFEM_OK;
}

otfem_status ensure_scratch(otfem_model *m, int64_t batch) {
  if (m->scratch_batch >= batch) return OTFEM_OK;
  free_p(m->Xt); free_p(m->H); free_p(m->Z); free_p(m->G); free_p(m->G4); free_p(m->Y);
  free_p(m->part); free_p(m->loss_part);
  m->scratch_batch = 0;
  const int64_t nparts = batch / 128;  // upper bound on workgroups of any kernel below
  MCHK(m, malloc_n(&m->Xt, 16 * batch));
  MCHK(m, malloc_n(&m->H, 3 * 128 * batch));
  MCHK(m, malloc_n(&m->Z, 3 * 128 * batch));
  MCHK(m, malloc_n(&m->G, 3 * 128 * batch));
  MCHK(m, malloc_n(&m->G4, 4 * batch));
  MCHK(m, malloc_n(&m->Y, 4 * batch));
  MCHK(m, malloc_n(&m->part, nparts * kMlpParams));
  MCHK(m, malloc_n(&m->loss_part, nparts));
  m->scratch_batch = batch;
  return OTFEM_OK;
}

int64_t chunk_for(int64_t batch) { return batch % 1024 == 0 ? 1024 : 128; }

FwdArgs fwd_args(otfem_model *m, const float *X, const float4 *T, int64_t count, int64_t batch,
                 int64_t nvalid) {
  FwdArgs a{};
  a.wf = reinterpret_cast<const bf16x8 *>(m->wf);
  a.theta = m->theta;
  a.X = X;
  a.T = T;
  a.count = count;
  a.batch = batch;
  a.nvalid = nvalid;
  a.Xt = m->Xt;
  a.Z = m->Z;
  a.H = m->H;
  a.Y = m->Y;
  a.G4 = m->G4;
  a.loss_part = m->loss_part;
  return a;
}

// Backward from dL/dy in G4 and the weight-gradient partials.
otfem_status run_bwd(otfem_model *m, int64_t batch, int *nparts) {
  BwdArgs b{};
  b.wb = reinterpret_cast<const bf16x8 *>(m->wb);
  b.batch = batch;
  b.Z = m->Z;
  b.G4 = m->G4;
  b.G = m->G;
  hipLaunchKernelGGL(k_mlp_bwd, dim3((unsigned)(batch / 128)), dim3(256), 0, m->stream, b);
  MCHK(m, hipGetLastError());
  GradArgs g{};
  g.G = m->G;
  g.G4 = m->G4;
  g.H = m->H;
  g.Xt = m->Xt;
  g.batch = batch;
  g.chunk = chunk_for(batch);
  g.part = m->part;
  const int64_t np = batch / g.chunk;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((unsigned)np), dim3(256), 0, m->stream, g);
  MCHK(m, hipGetLastError());
  *nparts = (int)np;
  return OTFEM_OK;
}

otfem_status run_adam(otfem_model *m, int nparts, double lr, int nloss, float *loss_out) {
  const float b1 = 0.9f, b2 = 0.999f;
  ++m->adam_t;
  AdamArgs a{};
  a.theta = m->theta;
  a.m = m->adam_m;
  a.v = m->adam_v;
  a.part = m->part;
  a.nparts = nparts;
  a.lr = (float)lr;
  a.beta1 = b1;
  a.beta2 = b2;
  a.eps = 1e-8f;
  a.corr1 = (float)(1.0 - std::pow((double)b1, (double)m->adam_t));
  a.corr2 = (float)(1.0 - std::pow((double)b2, (double)m->adam_t));
  a.fpos = m->fpos;
  a.bpos = m->bpos;
  a.wf = m->wf;
  a.wb = m->wb;
  a.loss_part = m->loss_part;
  a.nloss = nloss;
  a.loss_out = loss_out;
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream, a);
  MCHK(m, hipGetLastError());
  return OTFEM_OK;
}

otfem_status ensure_loss_hist(otfem_model *m, int steps) {
  if (m->loss_hist_cap >= steps) return OTFEM_OK;
  free_p(m->loss_hist);
  m->loss_hist_cap = 0;
  MCHK(m, malloc_n(&m->loss_hist, steps));
  m->loss_hist_cap = steps;
  return OTFEM_OK;
}

otfem_status model_device_check(otfem_model *m) {
  int dev = 0;
  MCHK(m, hipGetDevice(&dev));
  if (dev != m->device)
    return mfail(m, OTFEM_STATE, "the model was created on another device");
  return OTFEM_OK;
}

std::vector<float> rows16(const float *features, int64_t n) {
  std::vector<float> x(16 * n, 0.0f);
  for (int64_t c = 0; c < n; ++c)
    for (int k = 0; k < kMlpIn; ++k) x[16 * c + k] = features[kMlpIn * c + k];
  return x;
}

}  // namespace

namespace otfem {

bool model_view(otfem_model *m, ModelView *out) {
  if (!m || !out) return false;
  out->device = m->device;
  out->wf = m->wf;
  out->theta = m->theta;
  out->X = m->X;
  out->S = m->S;
  out->T = m->T;
  out->capacity = m->cap;
  out->count = m->count;
  return true;
}

void model_commit(otfem_model *m, int64_t n) {
  if (m) m->count += n;
}

int model_train_residual(otfem_model *m, const ResidualOp &op, int steps, int K, double lr,
                         uint64_t seed, float *loss, std::string *err) {
  return model_residual(m, op, steps, K, lr, seed, loss, nullptr, nullptr, nullptr, err);
}

int model_residual(otfem_model *m, const ResidualOp &op, int steps, int K, double lr,
                   uint64_t seed, float *loss, const int32_t *given_steps, float *grad,
                   double *loss_value, std::string *err) {
  auto bad = [&](otfem_status st, const std::string &msg) {
    m->err = msg;
    if (err) *err = msg;
    return (int)st;
  };
  if (given_steps) steps = 1;  // one loss / gradient evaluation, no update
  if (steps < 0 || K <= 0 || !(lr > 0)) return bad(OTFEM_INVALID_ARG, "steps >= 0, K > 0, lr > 0");
  const int64_t n = op.n;
  if (n <= 0 || m->count == 0 || m->count % n != 0)
    return bad(OTFEM_STATE, "the sample store does not hold whole steps of this mesh");
  if (steps == 0) return OTFEM_OK;
  otfem_status st = model_device_check(m);
  if (st) return bad(st, m->err);
  if (hipDeviceSynchronize() != hipSuccess) return bad(OTFEM_HIP_ERROR, "device synchronize");
  const int64_t n_steps = m->count / n;
  if (given_steps)
    for (int k = 0; k < K; ++k)
      if (given_steps[k] < 0 || given_steps[k] >= n_steps) return bad(OTFEM_INVALID_ARG, "step index");
  const int64_t nvalid = (int64_t)K * n, batch = (nvalid + 127) / 128 * 128;
  st = ensure_scratch(m, batch);
  if (!st) st = ensure_loss_hist(m, steps);
  if (st) return bad(st, m->err);
  const int64_t nnz = (int64_t)op.cols.size();
  DevBuf<int64_t> rp;
  DevBuf<int32_t> cols, dsteps;
  DevBuf<float> blocks;
  DevBuf<float4> E, Q, Gv;
  DevBuf<double> Q0;
  const int64_t rows = (nvalid + 255) / 256;
  if (malloc_n(&rp.p, n + 1) || malloc_n(&cols.p, nnz) || malloc_n(&blocks.p, 9 * nnz) ||
      malloc_n(&dsteps.p, K) || malloc_n(&E.p, nvalid) || malloc_n(&Q.p, nvalid) ||
      malloc_n(&Gv.p, nvalid) || malloc_n(&Q0.p, n_steps))
    return bad(OTFEM_HIP_ERROR, "hipMalloc of residual-training buffers");
  // the loss partials of k_res_grad need (batch + 255) / 256 <= batch / 128 slots
  hipStream_t s = m->stream;
  if (hipMemcpyAsync(rp.p, op.row_ptr.data(), (n + 1) * 8, hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(cols.p, op.cols.data(), nnz * 4, hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(blocks.p, op.blocks.data(), 9 * nnz * 4, hipMemcpyHostToDevice, s))
    return bad(OTFEM_HIP_ERROR, "upload of the residual operator");
  // |A e0|^2 of every recorded step (the plain start's residual)
  {
    DevBuf<float4> E0, Q0v;
    if (malloc_n(&E0.p, m->count) || malloc_n(&Q0v.p, m->count))
      return bad(OTFEM_HIP_ERROR, "hipMalloc");
    const unsigned g = (unsigned)((m->count + 255) / 256);
    hipLaunchKernelGGL(k_res_err, dim3(g), dim3(256), 0, s, nullptr, m->X, m->T, m->S, nullptr, n,
                       m->count, E0.p);
    hipLaunchKernelGGL(k_bsr_spmv, dim3(g), dim3(256), 0, s, rp.p, cols.p, blocks.p, n, n_steps,
                       E0.p, Q0v.p);
    hipLaunchKernelGGL(k_vec_norms, dim3((unsigned)n_steps), dim3(256), 0, s, Q0v.p, n, Q0.p);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      return bad(OTFEM_HIP_ERROR, "residual of the plain start");
  }
  if (given_steps &&
      hipMemcpyAsync(dsteps.p, given_steps, K * sizeof(int32_t), hipMemcpyHostToDevice, s))
    return bad(OTFEM_HIP_ERROR, "step upload");
  for (int t = 0; t < steps; ++t) {
    if (!given_steps)
      hipLaunchKernelGGL(k_pick_steps, dim3((K + 255) / 256), dim3(256), 0, s, dsteps.p, K,
                         n_steps, seed, m->adam_t);
    FwdArgs a = fwd_args(m, m->X, m->T, m->count, batch, nvalid);
    a.pick = PICK_STEPS;
    a.steps = dsteps.p;
    a.n_free = n;
    a.mse = 0;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)(batch / 128)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_res_err, dim3((unsigned)rows), dim3(256), 0, s, m->Y, m->X, m->T, m->S,
                       dsteps.p, n, nvalid, E.p);
    hipLaunchKernelGGL(k_bsr_spmv, dim3((unsigned)rows), dim3(256), 0, s, rp.p, cols.p, blocks.p, n,
                       (int64_t)K, E.p, Q.p);
    hipLaunchKernelGGL(k_bsr_spmv, dim3((unsigned)rows), dim3(256), 0, s, rp.p, cols.p, blocks.p, n,
                       (int64_t)K, Q.p, Gv.p);
    // G4 of the padded columns is written as well: the grid covers the batch
    const unsigned gb = (unsigned)((batch + 255) / 256);
    hipLaunchKernelGGL(k_res_grad, dim3(gb), dim3(256), 0, s, Q.p, Gv.p, Q0.p, dsteps.p, m->X,
                       m->S, n, K, nvalid, batch, m->G4, m->loss_part);
    if (hipGetLastError() != hipSuccess) return bad(OTFEM_HIP_ERROR, "residual-loss kernels");
    int nparts = 0;
    st = run_bwd(m, batch, &nparts);
    if (st) return bad(st, m->err);
    if (given_steps) {
      hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, m->part,
                         nparts, m->grad);
      std::vector<double> lp(gb);
      if (hipGetLastError() != hipSuccess ||
          (grad && hipMemcpyAsync(grad, m->grad, kMlpParams * 4, hipMemcpyDeviceToHost, s)) ||
          hipMemcpyAsync(lp.data(), m->loss_part, gb * 8, hipMemcpyDeviceToHost, s) ||
          hipStreamSynchronize(s))
        return bad(OTFEM_HIP_ERROR, "residual gradient");
      if (loss_value) {
        double v = 0.0;
        for (double x : lp) v += x;
        *loss_value = v;
      }
      return OTFEM_OK;
    }
    st = run_adam(m, nparts, lr, (int)gb, m->loss_hist + t);
    if (st) return bad(st, m->err);
  }
  if (loss && hipMemcpyAsync(loss, m->loss_hist, steps * sizeof(float), hipMemcpyDeviceToHost, s))
    return bad(OTFEM_HIP_ERROR, "loss copy");
  if (hipStreamSynchronize(s) != hipSuccess) return bad(OTFEM_HIP_ERROR, "training");
  return OTFEM_OK;
}

}  // namespace otfem

extern "C" {

otfem_status otfem_model_create(uint64_t seed, otfem_model_handle *out) {
  if (!out) return OTFEM_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return OTFEM_NO_DEVICE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return OTFEM_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return OTFEM_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return OTFEM_NO_DEVICE;
  otfem_model *m = new otfem_model();
  m->device = dev;
  build_positions(m->fpos_host, m->bpos_host);
  auto bail = [&](otfem_status st) {
    otfem_model_destroy(m);
    return st;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      malloc_n(&m->theta, kMlpParams) || malloc_n(&m->adam_m, kMlpParams) ||
      malloc_n(&m->adam_v, kMlpParams) || malloc_n(&m->wf, kFragFwd) ||
      malloc_n(&m->wb, kFragBwd) || malloc_n(&m->fpos, kMlpParams) ||
      malloc_n(&m->bpos, kMlpParams) || malloc_n(&m->wsum_part, kWsumBlocks) ||
      malloc_n(&m->grad, kMlpParams))
    return bail(OTFEM_HIP_ERROR);
  if (hipMemcpy(m->fpos, m->fpos_host.data(), kMlpParams * 4, hipMemcpyHostToDevice) ||
      hipMemcpy(m->bpos, m->bpos_host.data(), kMlpParams * 4, hipMemcpyHostToDevice))
    return bail(OTFEM_HIP_ERROR);
  // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for the hidden layers; the
  // output layer starts at y = (2, -1, 0): quadratic extrapolation.
  std::mt19937_64 rng(seed);
  std::vector<float> th(kMlpParams, 0.0f);
  auto fill = [&](int off, int n, int fan_in) {
    std::uniform_real_distribution<float> U(-1.0f / std::sqrt((float)fan_in),
                                            1.0f / std::sqrt((float)fan_in));
    for (int k = 0; k < n; ++k) th[off + k] = U(rng);
  };
  fill(kOffW1, 128 * kMlpIn, kMlpIn);
  fill(kOffB1, 128, kMlpIn);
  fill(kOffW2, 128 * 128, 128);
  fill(kOffB2, 128, 128);
  fill(kOffW3, 128 * 128, 128);
  fill(kOffB3, 128, 128);
  th[kOffB4] = 2.0f;
  th[kOffB4 + 1] = -1.0f;
  if (upload_params(m, th) != OTFEM_OK) return bail(OTFEM_HIP_ERROR);
  *out = m;
  return OTFEM_OK;
}

otfem_status otfem_model_destroy(otfem_model_handle m) {
  if (!m) return OTFEM_INVALID_ARG;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  free_p(m->theta); free_p(m->adam_m); free_p(m->adam_v); free_p(m->wf); free_p(m->wb);
  free_p(m->fpos); free_p(m->bpos); free_p(m->X); free_p(m->S); free_p(m->T);
  free_p(m->Xt); free_p(m->H); free_p(m->Z); free_p(m->G); free_p(m->G4); free_p(m->Y);
  free_p(m->part); free_p(m->grad); free_p(m->loss_part); free_p(m->loss_hist);
  free_p(m->wsum_part);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return OTFEM_OK;
}

const char *otfem_model_last_error(otfem_model_handle m) { return m ? m->err.c_str() : "null model"; }

otfem_status otfem_model_get_params(otfem_model_handle m, float *theta) {
  if (!m || !theta) return OTFEM_INVALID_ARG;
  MCHK(m, hipDeviceSynchronize());
  MCHK(m, hipMemcpy(theta, m->theta, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  return OTFEM_OK;
}

// File: "OTFEMML1", uint32 version 1, uint32 parameter count, int64 Adam
// step, then theta, Adam first and second moments (FP32 each).
otfem_status otfem_model_save(otfem_model_handle m, const char *path) {
  if (!m || !path) return OTFEM_INVALID_ARG;
  std::vector<float> th(kMlpParams), am(kMlpParams), av(kMlpParams);
  MCHK(m, hipDeviceSynchronize());
  MCHK(m, hipMemcpy(th.data(), m->theta, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  MCHK(m, hipMemcpy(am.data(), m->adam_m, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  MCHK(m, hipMemcpy(av.data(), m->adam_v, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  std::ofstream f(path, std::ios::binary);
  if (!f) return mfail(m, OTFEM_INVALID_ARG, std::string("cannot write ") + path);
  const uint32_t version = 1, count = kMlpParams;
  const int64_t t = m->adam_t;
  f.write("OTFEMML1", 8);
  f.write(reinterpret_cast<const char *>(&version), 4);
  f.write(reinterpret_cast<const char *>(&count), 4);
  f.write(reinterpret_cast<const char *>(&t), 8);
  for (auto *v : {&th, &am, &av}) f.write(reinterpret_cast<const char *>(v->data()), kMlpParams * 4);
  if (!f) return mfail(m, OTFEM_INVALID_ARG, std::string("write failed: ") + path);
  return OTFEM_OK;
}

otfem_status otfem_model_load(otfem_model_handle m, const char *path) {
  if (!m || !path) return OTFEM_INVALID_ARG;
  otfem_status st = model_device_check(m);
  if (st) return st;
  std::ifstream f(path, std::ios::binary);
  if (!f) return mfail(m, OTFEM_INVALID_ARG, std::string("cannot read ") + path);
  char magic[8];
  uint32_t version = 0, count = 0;
  int64_t t = 0;
  f.read(magic, 8);
  f.read(reinterpret_cast<char *>(&version), 4);
  f.read(reinterpret_cast<char *>(&count), 4);
  f.read(reinterpret_cast<char *>(&t), 8);
  if (!f || std::memcmp(magic, "OTFEMML1", 8) != 0 || version != 1 || count != (uint32_t)kMlpParams ||
      t < 0)
    return mfail(m, OTFEM_INVALID_ARG, std::string("not a model file of this network: ") + path);
  std::vector<float> th(kMlpParams), am(kMlpParams), av(kMlpParams);
  for (auto *v : {&th, &am, &av}) f.read(reinterpret_cast<char *>(v->data()), kMlpParams * 4);
  if (!f) return mfail(m, OTFEM_INVALID_ARG, std::string("model file is truncated: ") + path);
  MCHK(m, hipDeviceSynchronize());
  st = upload_params(m, th);  // also refreshes the BF16 fragments; zeroes Adam
  if (st) return st;
  MCHK(m, hipMemcpy(m->adam_m, am.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->adam_v, av.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  m->adam_t = t;
  return OTFEM_OK;
}

otfem_status otfem_model_set_params(otfem_model_handle m, const float *theta) {
  if (!m || !theta) return OTFEM_INVALID_ARG;
  otfem_status st = model_device_check(m);
  if (st) return st;
  MCHK(m, hipDeviceSynchronize());
  return upload_params(m, std::vector<float>(theta, theta + kMlpParams));
}

otfem_status otfem_model_forward(otfem_model_handle m, const float *features, int64_t n,
                                 float *out) {
  if (!m || n < 0 || (n > 0 && (!features || !out))) return OTFEM_INVALID_ARG;
  if (n == 0) return OTFEM_OK;
  otfem_status st = model_device_check(m);
  if (st) return st;
  const std::vector<float> xr = rows16(features, n);
  DevBuf<float> dX, dY;
  MCHK(m, malloc_n(&dX.p, 16 * n));
  MCHK(m, malloc_n(&dY.p, 3 * n));
  MCHK(m, hipMemcpyAsync(dX.p, xr.data(), xr.size() * 4, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_mlp_rows, dim3((unsigned)((n + 127) / 128)), dim3(256), 0, m->stream,
                     reinterpret_cast<const bf16x8 *>(m->wf), m->theta, dX.p, n, dY.p);
  MCHK(m, hipGetLastError());
  MCHK(m, hipMemcpyAsync(out, dY.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return OTFEM_OK;
}

otfem_status otfem_model_gradient(otfem_model_handle m, const float *features,
                                  const float *targets, int64_t n, double *loss, float *grad) {
  if (!m || n <= 0 || !features || !targets) return OTFEM_INVALID_ARG;
  otfem_status st = model_device_check(m);
  if (st) return st;
  double wsum = 0.0;
  for (int64_t c = 0; c < n; ++c) {
    if (!(targets[4 * c + 3] >= 0)) return mfail(m, OTFEM_INVALID_ARG, "negative loss weight");
    wsum += targets[4 * c + 3];
  }
  if (!(wsum > 0)) return mfail(m, OTFEM_INVALID_ARG, "all loss weights are zero");
  const int64_t batch = (n + 127) / 128 * 128;
  st = ensure_scratch(m, batch);
  if (st) return st;
  const std::vector<float> xr = rows16(features, n);
  DevBuf<float> dX;
  DevBuf<float4> dT;
  MCHK(m, malloc_n(&dX.p, 16 * n));
  MCHK(m, malloc_n(&dT.p, n));
  MCHK(m, hipMemcpyAsync(dX.p, xr.data(), xr.size() * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(dT.p, targets, n * sizeof(float4), hipMemcpyHostToDevice, m->stream));
  FwdArgs a = fwd_args(m, dX.p, dT.p, n, batch, n);
  a.pick = PICK_CONTIGUOUS;
  a.base = 0;
  a.mse = 1;
  a.inv_w = (float)((double)n / wsum);
  hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)(batch / 128)), dim3(256), 0, m->stream, a);
  MCHK(m, hipGetLastError());
  int nparts = 0;
  st = run_bwd(m, batch, &nparts);
  if (st) return st;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream,
                     m->part, nparts, m->grad);
  MCHK(m, hipGetLastError());
  std::vector<double> lp(batch / 128);
  if (grad)
    MCHK(m, hipMemcpyAsync(grad, m->grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost,
                           m->stream));
  MCHK(m, hipMemcpyAsync(lp.data(), m->loss_part, lp.size() * sizeof(double),
                         hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (loss) {
    double s = 0.0;
    for (double v : lp) s += v;
    *loss = s;
  }
  return OTFEM_OK;
}

otfem_status otfem_model_reserve(otfem_model_handle m, int64_t capacity) {
  if (!m || capacity < 0) return OTFEM_INVALID_ARG;
  if (capacity <= m->cap) return OTFEM_OK;
  otfem_status st = model_device_check(m);
  if (st) return st;
  MCHK(m, hipDeviceSynchronize());
  float *X = nullptr, *S = nullptr;
  float4 *T = nullptr;
  if (malloc_n(&X, 16 * capacity) || malloc_n(&S, capacity) || malloc_n(&T, capacity)) {
    free_p(X);
    free_p(S);
    free_p(T);
    return mfail(m, OTFEM_HIP_ERROR, "hipMalloc of the sample store");
  }
  if (m->count) {
    MCHK(m, hipMemcpy(X, m->X, 16 * m->count * sizeof(float), hipMemcpyDeviceToDevice));
    MCHK(m, hipMemcpy(S, m->S, m->count * sizeof(float), hipMemcpyDeviceToDevice));
    MCHK(m, hipMemcpy(T, m->T, m->count * sizeof(float4), hipMemcpyDeviceToDevice));
  }
  free_p(m->X);
  free_p(m->S);
  free_p(m->T);
  m->X = X;
  m->S = S;
  m->T = T;
  m->cap = capacity;
  return OTFEM_OK;
}

otfem_status otfem_model_samples(otfem_model_handle m, int64_t *count, int64_t *capacity) {
  if (!m) return OTFEM_INVALID_ARG;
  if (count) *count = m->count;
  if (capacity) *capacity = m->cap;
  return OTFEM_OK;
}

otfem_status otfem_model_clear_samples(otfem_model_handle m) {
  if (!m) return OTFEM_INVALID_ARG;
  MCHK(m, hipDeviceSynchronize());
  m->count = 0;
  return OTFEM_OK;
}

otfem_status otfem_model_get_samples(otfem_model_handle m, int64_t offset, int64_t n,
                                     float *features, float *targets) {
  if (!m || offset < 0 || n < 0 || offset + n > m->count) return OTFEM_INVALID_ARG;
  if (n == 0) return OTFEM_OK;
  MCHK(m, hipDeviceSynchronize());
  if (features) {
    std::vector<float> xr(16 * n);
    MCHK(m, hipMemcpy(xr.data(), m->X + 16 * offset, xr.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < n; ++c)
      for (int k = 0; k < kMlpIn; ++k) features[kMlpIn * c + k] = xr[16 * c + k];
  }
  if (targets)
    MCHK(m, hipMemcpy(targets, m->T + offset, n * sizeof(float4), hipMemcpyDeviceToHost));
  return OTFEM_OK;
}

otfem_status otfem_model_add_samples(otfem_model_handle m, const float *features,
                                     const float *targets, int64_t n) {
  if (!m || n < 0 || (n > 0 && (!features || !targets))) return OTFEM_INVALID_ARG;
  if (m->count + n > m->cap) return mfail(m, OTFEM_STATE, "sample store full (otfem_model_reserve)");
  if (n == 0) return OTFEM_OK;
  for (int64_t c = 0; c < n; ++c)
    if (!(targets[4 * c + 3] >= 0)) return mfail(m, OTFEM_INVALID_ARG, "negative loss weight");
  MCHK(m, hipDeviceSynchronize());
  const std::vector<float> xr = rows16(features, n);
  MCHK(m, hipMemcpy(m->X + 16 * m->count, xr.data(), xr.size() * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->T + m->count, targets, n * sizeof(float4), hipMemcpyHostToDevice));
  MCHK(m, hipMemset(m->S + m->count, 0, n * sizeof(float)));
  m->count += n;
  return OTFEM_OK;
}

otfem_status otfem_model_train(otfem_model_handle m, int steps, int batch, double lr,
                               uint64_t seed, float *loss) {
  if (!m || steps < 0 || batch <= 0 || batch % 128 != 0 || !(lr > 0))
    return mfail(m, OTFEM_INVALID_ARG, "steps >= 0, batch a positive multiple of 128, lr > 0");
  if (m->count == 0) return mfail(m, OTFEM_STATE, "no training samples recorded");
  if (steps == 0) return OTFEM_OK;
  otfem_status st = model_device_check(m);
  if (st) return st;
  MCHK(m, hipDeviceSynchronize());  // samples recorded by solver streams
  st = ensure_scratch(m, batch);
  if (!st) st = ensure_loss_hist(m, steps);
  if (st) return st;
  hipLaunchKernelGGL(k_weight_sum, dim3(kWsumBlocks), dim3(256), 0, m->stream, m->T, m->count,
                     m->wsum_part);
  MCHK(m, hipGetLastError());
  std::vector<double> wp(kWsumBlocks);
  MCHK(m, hipMemcpyAsync(wp.data(), m->wsum_part, kWsumBlocks * sizeof(double),
                         hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  double wsum = 0.0;
  for (double v : wp) wsum += v;
  if (!(wsum > 0)) return mfail(m, OTFEM_STATE, "all recorded samples have zero weight");
  for (int k = 0; k < steps; ++k) {
    FwdArgs a = fwd_args(m, m->X, m->T, m->count, batch, batch);
    a.pick = PICK_RANDOM;
    a.seed = seed;
    a.step = m->adam_t;
    a.mse = 1;
    a.inv_w = (float)((double)m->count / wsum);
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)(batch / 128)), dim3(256), 0, m->stream, a);
    MCHK(m, hipGetLastError());
    int nparts = 0;
    st = run_bwd(m, batch, &nparts);
    if (!st) st = run_adam(m, nparts, lr, batch / 128, m->loss_hist + k);
    if (st) return st;
  }
  if (loss)
    MCHK(m, hipMemcpyAsync(loss, m->loss_hist, steps * sizeof(float), hipMemcpyDeviceToHost,
                           m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return OTFEM_OK;
}

}  // extern "C"
