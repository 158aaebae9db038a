// bl_mlp.hip -- the learned node relaxation: a 26-64-64-2 multilayer
// perceptron that proposes each relaxed node's displacement from its
// 8-node stencil, trained on pairs the plain relaxation produces.  This
// part has no counterpart in the reference.
//
// Features of node n (with its stencil table row nb[0..7], §bl_relax.hip):
// with h^2 the mean area of the four elements around n,
//   f[0..15]  neighbour offsets (x_nb - x_n, y_nb - y_n)/h, table order
//   f[16..19] the four element areas / h^2 (counter-clockwise from +k)
//   f[20..23] the corner Jacobians at n of those elements / h^2
//   f[24..25] the mean neighbour offset / h
// Target: the plain relaxation's displacement (x_R - x, y_R - y)/h.
//
// Inference is one kernel: each wave takes 16 nodes, computes their
// features and runs the three layers on the matrix cores
// (v_mfma_f32_16x16x16_bf16: BF16 weights and activations, FP32
// accumulation), with the activations going through LDS between layers.
// Training is one thread per sample in FP32 with the forward pass rounding
// weights and activations to BF16 as the matrix cores see them
// (straight-through gradients); per-workgroup gradients are summed in a
// fixed order, so training is reproducible run to run; Adam updates FP32
// master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "bl_internal.h"
#include "bl_math.h"

namespace bl {

constexpr int kIn = 26, kInPad = 32, kHid = 64, kOut = 2;
// parameter layout (FP32 master): W1 (64,26) row-major, b1, W2 (64,64), b2,
// W3 (2,64), b3
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid,
              kOffB2 = kOffW2 + kHid * kHid, kOffW3 = kOffB2 + kHid,
              kOffB3 = kOffW3 + kOut * kHid, kNParam = kOffB3 + kOut;
constexpr int kTrainBlock = 64;  // samples per training workgroup

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even BF16 and back
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ inline float bfr(float f) { return bf2f(f2bf(f)); }

// the features of a relaxed node (fp64 arithmetic, FP32 result); h the
// length scale
__device__ __forceinline__ void node_features(const int32_t* __restrict__ ring,
                                              const double* __restrict__ x,
                                              const double* __restrict__ y, int n,
                                              float f[kIn], double& h) {
  const int4 qa = reinterpret_cast<const int4*>(ring)[2 * n];
  const int4 qb = reinterpret_cast<const int4*>(ring)[2 * n + 1];
  const int nb[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
  const double xn = x[n], yn = y[n];
  double dx[8], dy[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) { dx[i] = x[nb[i]] - xn; dy[i] = y[nb[i]] - yn; }
  // counter-clockwise ring P0..P7 (P0 = +k) from the table order
  const int ccw[8] = {4, 2, 1, 0, 3, 5, 6, 7};
  double area[4], jac[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int a = ccw[2 * q], b = ccw[2 * q + 1], c = ccw[(2 * q + 2) & 7];
    area[q] = 0.5 * (dx[b] * (dy[c] - dy[a]) - (dx[c] - dx[a]) * dy[b]);
    jac[q] = dx[a] * dy[c] - dx[c] * dy[a];
  }
  const double h2 = (area[0] + area[1] + area[2] + area[3]) / 4.0;
  h = sqrt(h2 > 1.0e-300 ? h2 : 1.0e-300);
  const double ih = 1.0 / h, ih2 = 1.0 / (h * h);
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    f[2 * i] = (float)(dx[i] * ih);
    f[2 * i + 1] = (float)(dy[i] * ih);
    sx = sx + dx[i];
    sy = sy + dy[i];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    f[16 + q] = (float)(area[q] * ih2);
    f[20 + q] = (float)(jac[q] * ih2);
  }
  f[24] = (float)(sx / 8.0 * ih);
  f[25] = (float)(sy / 8.0 * ih);
}

// BF16 weights packed as the matrix cores' B operand, one 8-byte group per
// (layer tile, k-step, lane): lane l of tile t, k-step s holds
// W[16t + l%16][16s + 4(l/16) + j], j = 0..3 (zero past the layer's width)
struct MlpPacked {
  const uint64_t* w1;  // (4 tiles, 2 k-steps, 64 lanes)
  const uint64_t* w2;  // (4, 4, 64)
  const uint64_t* w3;  // (1, 4, 64), rows 0,1 real
  const float* b1;
  const float* b2;
  const float* b3;
};

__device__ __forceinline__ bf16x4 as_bf16x4(uint64_t v) {
  bf16x4 r;
  r[0] = (short)(v & 0xffff);
  r[1] = (short)((v >> 16) & 0xffff);
  r[2] = (short)((v >> 32) & 0xffff);
  r[3] = (short)((v >> 48) & 0xffff);
  return r;
}
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
  return r;
}

// the three layers for the 16 rows of this wave: in0, in1 the A operands of
// the input (k-steps 0, 1), lds a 16 x 64 BF16 scratch of the wave; y out
// in the D layout: lane l holds rows 4(l/16)+j of column l%16.  Every wave
// of the workgroup calls it (the barriers are workgroup-wide).
__device__ __forceinline__ f32x4 mlp16(const MlpPacked& W, const bf16x4 in0, const bf16x4 in1,
                                       uint16_t* lds) {
  const int l = threadIdx.x & 63, col = l & 15, grp = l >> 4;
  // layer 1
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(in0, as_bf16x4(W.w1[(t * 2 + 0) * 64 + l]),
                                                    acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(in1, as_bf16x4(W.w1[(t * 2 + 1) * 64 + l]),
                                                    acc, 0, 0, 0);
    const float b = W.b1[16 * t + col];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float z = acc[j] + b;
      lds[(4 * grp + j) * kHid + 16 * t + col] = f2bf(z > 0.0f ? z : 0.0f);
    }
  }
  __syncthreads();
  bf16x4 a[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const uint16_t* p = lds + col * kHid + 16 * s + 4 * grp;
    a[s] = bf16x4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
  }
  __syncthreads();
  // layer 2
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 4; ++s)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[s], as_bf16x4(W.w2[(t * 4 + s) * 64 + l]),
                                                      acc, 0, 0, 0);
    const float b = W.b2[16 * t + col];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float z = acc[j] + b;
      lds[(4 * grp + j) * kHid + 16 * t + col] = f2bf(z > 0.0f ? z : 0.0f);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const uint16_t* p = lds + col * kHid + 16 * s + 4 * grp;
    a[s] = bf16x4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
  }
  // layer 3
  f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < 4; ++s)
    y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[s], as_bf16x4(W.w3[s * 64 + l]), y, 0, 0, 0);
  const float b = col < kOut ? W.b3[col] : 0.0f;
#pragma unroll
  for (int j = 0; j < 4; ++j) y[j] = y[j] + b;
  return y;
}

// inference on a feature buffer (n x 26 FP32) into y (n x 2)
__global__ void __launch_bounds__(256) k_mlp_forward(MlpPacked W, const float* __restrict__ xin,
                                                     int n, float* __restrict__ yout) {
  __shared__ uint16_t lds[4][16 * kHid];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 15, grp = l >> 4;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  const int r = row0 + col;
  float f[kInPad];
#pragma unroll
  for (int k = 0; k < kInPad; ++k) f[k] = (r < n && k < kIn) ? xin[(size_t)r * kIn + k] : 0.0f;
  const bf16x4 in0 = pack4(f[4 * grp], f[4 * grp + 1], f[4 * grp + 2], f[4 * grp + 3]);
  const bf16x4 in1 = pack4(f[16 + 4 * grp], f[17 + 4 * grp], f[18 + 4 * grp], f[19 + 4 * grp]);
  const f32x4 y = mlp16(W, in0, in1, lds[wave]);
  if (col < kOut) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rr = row0 + 4 * grp + j;
      if (rr < n) yout[(size_t)rr * kOut + col] = y[j];
    }
  }
}

// inference on the relaxed nodes of a mesh: the predicted displacement / h
// in pred (2, nnd) and h; gated like the ALE kernels
__global__ void __launch_bounds__(256) k_mlp_nodes(MlpPacked W, const int32_t* __restrict__ ring,
                                                   const int32_t* __restrict__ rlist, int nr,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y, int nnd,
                                                   float* __restrict__ pred,
                                                   double* __restrict__ hout,
                                                   int32_t* __restrict__ flag,
                                                   const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 0;
  __shared__ uint16_t lds[4][16 * kHid];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 15, grp = l >> 4;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  const int r = row0 + col;
  float f[kInPad];
#pragma unroll
  for (int k = 0; k < kInPad; ++k) f[k] = 0.0f;
  double h = 1.0;
  const int node = r < nr ? rlist[r] : -1;
  if (node >= 0) node_features(ring, x, y, node, f, h);
  if (node >= 0 && grp == 0) hout[node] = h;
  const bf16x4 in0 = pack4(f[4 * grp], f[4 * grp + 1], f[4 * grp + 2], f[4 * grp + 3]);
  const bf16x4 in1 = pack4(f[16 + 4 * grp], f[17 + 4 * grp], f[18 + 4 * grp], f[19 + 4 * grp]);
  const f32x4 yv = mlp16(W, in0, in1, lds[wave]);
  if (col < kOut) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rr = row0 + 4 * grp + j;
      if (rr < nr) pred[(size_t)col * nnd + rlist[rr]] = yv[j];
    }
  }
}

// the learned motion proposes x + h * pred for every node: pred and h are
// zero for nodes that are not relaxed (set when the network is attached),
// so those stay at x exactly
__device__ __forceinline__ double2 proposed(int n, const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const float* __restrict__ pred,
                                            const double* __restrict__ h, int nnd) {
  const double hn = h[n];
  return make_double2(x[n] + hn * (double)pred[n], y[n] + hn * (double)pred[(size_t)nnd + n]);
}

// the quality check of the proposed mesh: any element with a corner
// Jacobian <= 0 or a longest/shortest side ratio > 10 sets the fallback
__global__ void k_learned_check(int nel, int nnd, const int32_t* __restrict__ ielnd,
                                const double* __restrict__ x, const double* __restrict__ y,
                                const float* __restrict__ pred, const double* __restrict__ h,
                                int32_t* __restrict__ flag, const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  const int e = xcd_index();
  if (e >= nel) return;
  const int4 q = reinterpret_cast<const int4*>(ielnd)[e];
  const int nd[4] = {q.x - 1, q.y - 1, q.z - 1, q.w - 1};
  double px[4], py[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double2 p = proposed(nd[c], x, y, pred, h, nnd);
    px[c] = p.x;
    py[c] = p.y;
  }
  bool bad = false;
  double lmin = DBL_MAX, lmax = 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int a = (c + 3) & 3, b = (c + 1) & 3;
    const double j = (px[b] - px[c]) * (py[a] - py[c]) - (px[a] - px[c]) * (py[b] - py[c]);
    bad = bad || !(j > 0.0);
    const double ex = px[b] - px[c], ey = py[b] - py[c];
    const double len2 = ex * ex + ey * ey;
    lmin = len2 < lmin ? len2 : lmin;
    lmax = len2 > lmax ? len2 : lmax;
  }
  bad = bad || !(lmax <= 100.0 * lmin);
  if (bad) flag[0] = 1;
}

// the learned mesh velocity (x_proposed - x)/dt when the check passed; the
// step and fallback counters
__global__ void k_learned_velocity(int nnd, const double* __restrict__ x,
                                   const double* __restrict__ y, const float* __restrict__ pred,
                                   const double* __restrict__ h, double* __restrict__ wu,
                                   double* __restrict__ wv, double dt,
                                   int32_t* __restrict__ flag, const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  if (ts) dt = ts->dt;
  const int n = xcd_index();
  const int fell_back = flag[0];
  if (n == 0) {
    flag[1] = flag[1] + 1;
    if (fell_back) flag[2] = flag[2] + 1;
  }
  if (n >= nnd || fell_back) return;
  const double2 p = proposed(n, x, y, pred, h, nnd);
  wu[n] = (p.x - x[n]) / dt;
  wv[n] = (p.y - y[n]) / dt;
}

// training pairs from the plain relaxation's mesh velocity: features at x,
// target w * dt / h; every stride-th relaxed node, starting at the step
// number modulo stride, appended at pairs[count] while they fit
__device__ __forceinline__ int collect_phase(int stride, const DevTimestep* ts) {
  return ts ? ts->nstep % stride : 0;
}

__global__ void k_collect(const int32_t* __restrict__ ring, const int32_t* __restrict__ rlist,
                          int nr, int stride, const double* __restrict__ x,
                          const double* __restrict__ y, const double* __restrict__ wu,
                          const double* __restrict__ wv, double dt,
                          const int32_t* __restrict__ count, int cap, float* __restrict__ feat,
                          float* __restrict__ targ, const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  if (ts) dt = ts->dt;
  const int phase = collect_phase(stride, ts);
  const int take = nr > phase ? (nr - phase + stride - 1) / stride : 0;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int base = *count;
  if (i >= take || base + take > cap) return;
  const int n = rlist[phase + i * stride];
  float f[kIn];
  double h;
  node_features(ring, x, y, n, f, h);
  float* fo = feat + (size_t)(base + i) * kIn;
#pragma unroll
  for (int k = 0; k < kIn; ++k) fo[k] = f[k];
  targ[(size_t)(base + i) * kOut] = (float)(wu[n] * dt / h);
  targ[(size_t)(base + i) * kOut + 1] = (float)(wv[n] * dt / h);
}

__global__ void k_collect_end(int32_t* __restrict__ count, int nr, int stride, int cap,
                              const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  const int phase = collect_phase(stride, ts);
  const int take = nr > phase ? (nr - phase + stride - 1) / stride : 0;
  if (threadIdx.x == 0 && *count + take <= cap) *count = *count + take;
}

// ---- training --------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// one workgroup of kTrainBlock samples, a thread per sample: forward with
// BF16-rounded weights, inputs and hidden activations and FP32 sums;
// backward; this workgroup's gradient of the mean squared error over the
// batch (each parameter summed over the samples in order) into
// part[blockIdx][param], its loss sum into lpart[blockIdx].  Samples come
// from idx (explicit) or are drawn from npairs by (seed, it).  Per-sample
// vectors live in LDS rows padded to an odd stride (no bank conflicts).
constexpr int kInP = kIn + 1, kHidP = kHid + 1;

__global__ void __launch_bounds__(kTrainBlock)
k_mlp_train(const float* __restrict__ w, const float* __restrict__ feat,
            const float* __restrict__ targ, const int32_t* __restrict__ idx, int npairs,
            int batch, uint64_t seed, int it, float* __restrict__ part,
            float* __restrict__ lpart) {
  __shared__ float sw[kNParam];
  __shared__ float sx[kTrainBlock][kInP];
  __shared__ float sa1[kTrainBlock][kHidP];   // r(z1) in BF16
  __shared__ float sa2[kTrainBlock][kHidP];   // r(z2) in BF16
  __shared__ float sd1[kTrainBlock][kHidP];   // z1, then dL/dz1
  __shared__ float sd2[kTrainBlock][kHidP];   // z2, then dL/dz2
  __shared__ float sd3[kTrainBlock][kOut];    // dL/dy
  __shared__ float sloss[kTrainBlock];
  const int t = threadIdx.x;
  for (int p = t; p < kNParam; p += kTrainBlock) {
    const bool is_w = (p < kOffB1) || (p >= kOffW2 && p < kOffB2) || (p >= kOffW3 && p < kOffB3);
    sw[p] = is_w ? bfr(w[p]) : w[p];
  }
  const int s = blockIdx.x * kTrainBlock + t;
  const bool live = s < batch;
  int src = 0;
  if (live) {
    if (idx) src = idx[s];
    else src = (int)(mix64(seed ^ mix64(((uint64_t)it << 32) | (uint32_t)s)) % (uint64_t)npairs);
  }
  for (int k = 0; k < kIn; ++k) sx[t][k] = live ? bfr(feat[(size_t)src * kIn + k]) : 0.0f;
  const float t0 = live ? targ[(size_t)src * kOut] : 0.0f;
  const float t1 = live ? targ[(size_t)src * kOut + 1] : 0.0f;
  __syncthreads();
  for (int i = 0; i < kHid; ++i) {
    float acc = 0.0f;
    for (int k = 0; k < kIn; ++k) acc = fmaf(sw[kOffW1 + i * kIn + k], sx[t][k], acc);
    const float z = acc + sw[kOffB1 + i];
    sd1[t][i] = z;
    sa1[t][i] = bfr(z > 0.0f ? z : 0.0f);
  }
  for (int i = 0; i < kHid; ++i) {
    float acc = 0.0f;
    for (int k = 0; k < kHid; ++k) acc = fmaf(sw[kOffW2 + i * kHid + k], sa1[t][k], acc);
    const float z = acc + sw[kOffB2 + i];
    sd2[t][i] = z;
    sa2[t][i] = bfr(z > 0.0f ? z : 0.0f);
  }
  float loss = 0.0f;
  for (int o = 0; o < kOut; ++o) {
    float acc = 0.0f;
    for (int k = 0; k < kHid; ++k) acc = fmaf(sw[kOffW3 + o * kHid + k], sa2[t][k], acc);
    const float r = acc + sw[kOffB3 + o] - (o == 0 ? t0 : t1);
    loss = loss + r * r;
    // d(mean over batch and outputs of r^2)/dy
    sd3[t][o] = live ? 2.0f * r / (float)(batch * kOut) : 0.0f;
  }
  sloss[t] = live ? loss : 0.0f;
  for (int i = 0; i < kHid; ++i) {
    const float acc = fmaf(sw[kOffW3 + i], sd3[t][0], sw[kOffW3 + kHid + i] * sd3[t][1]);
    sd2[t][i] = sd2[t][i] > 0.0f ? acc : 0.0f;
  }
  for (int i = 0; i < kHid; ++i) {
    float acc = 0.0f;
    for (int k = 0; k < kHid; ++k) acc = fmaf(sw[kOffW2 + k * kHid + i], sd2[t][k], acc);
    sd1[t][i] = sd1[t][i] > 0.0f ? acc : 0.0f;
  }
  __syncthreads();
  // this workgroup's gradient: parameter p summed over the samples in order
  float* out = part + (size_t)blockIdx.x * kNParam;
  for (int p = t; p < kNParam; p += kTrainBlock) {
    float g = 0.0f;
    if (p < kOffB1) {
      const int i = p / kIn, k = p % kIn;
      for (int q = 0; q < kTrainBlock; ++q) g = fmaf(sd1[q][i], sx[q][k], g);
    } else if (p < kOffW2) {
      const int i = p - kOffB1;
      for (int q = 0; q < kTrainBlock; ++q) g = g + sd1[q][i];
    } else if (p < kOffB2) {
      const int i = (p - kOffW2) / kHid, k = (p - kOffW2) % kHid;
      for (int q = 0; q < kTrainBlock; ++q) g = fmaf(sd2[q][i], sa1[q][k], g);
    } else if (p < kOffW3) {
      const int i = p - kOffB2;
      for (int q = 0; q < kTrainBlock; ++q) g = g + sd2[q][i];
    } else if (p < kOffB3) {
      const int o = (p - kOffW3) / kHid, k = (p - kOffW3) % kHid;
      for (int q = 0; q < kTrainBlock; ++q) g = fmaf(sd3[q][o], sa2[q][k], g);
    } else {
      const int o = p - kOffB3;
      for (int q = 0; q < kTrainBlock; ++q) g = g + sd3[q][o];
    }
    out[p] = g;
  }
  if (t == 0) {
    float l = 0.0f;
    for (int q = 0; q < kTrainBlock; ++q) l = l + sloss[q];
    lpart[blockIdx.x] = l;
  }
}

// the gradient: workgroup partials summed in order; loss likewise
__global__ void k_mlp_reduce(int nblk, const float* __restrict__ part,
                             const float* __restrict__ lpart, float* __restrict__ grad,
                             float* __restrict__ loss, int batch) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNParam) {
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g = g + part[(size_t)b * kNParam + p];
    grad[p] = g;
  }
  if (p == 0) {
    float l = 0.0f;
    for (int b = 0; b < nblk; ++b) l = l + lpart[b];
    *loss = l / (float)(batch * kOut);
  }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master weights
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const float b1 = 0.9f, b2 = 0.999f;
  m[p] = b1 * m[p] + (1.0f - b1) * g[p];
  v[p] = b2 * v[p] + (1.0f - b2) * g[p] * g[p];
  const float mh = m[p] / (1.0f - powf(b1, (float)step));
  const float vh = v[p] / (1.0f - powf(b2, (float)step));
  w[p] = w[p] - lr * mh / (sqrtf(vh) + 1.0e-8f);
}

// pack the master weights as BF16 B operands and FP32 biases
__global__ void k_mlp_pack(const float* __restrict__ w, uint64_t* __restrict__ w1p,
                           uint64_t* __restrict__ w2p, uint64_t* __restrict__ w3p,
                           float* __restrict__ bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  auto pack = [&](int rows, int cols, int off, int t, int s, int l) {
    const int row = 16 * t + (l & 15);
    uint64_t v = 0;
    for (int j = 0; j < 4; ++j) {
      const int c = 16 * s + 4 * (l >> 4) + j;
      const float f = (row < rows && c < cols) ? w[off + row * cols + c] : 0.0f;
      v |= (uint64_t)f2bf(f) << (16 * j);
    }
    return v;
  };
  if (i < 4 * 2 * 64) {
    const int t = i / 128, s = (i / 64) % 2, l = i % 64;
    w1p[i] = pack(kHid, kIn, kOffW1, t, s, l);
  }
  if (i < 4 * 4 * 64) {
    const int t = i / 256, s = (i / 64) % 4, l = i % 64;
    w2p[i] = pack(kHid, kHid, kOffW2, t, s, l);
  }
  if (i < 4 * 64) {
    const int s = i / 64, l = i % 64;
    w3p[i] = pack(kOut, kHid, kOffW3, 0, s, l);
  }
  if (i < kHid) { bias[i] = w[kOffB1 + i]; bias[kHid + i] = w[kOffB2 + i]; }
  if (i < kOut) bias[2 * kHid + i] = w[kOffB3 + i];
}

}  // namespace bl

using namespace bl;

struct bl_mlp {
  float* w = nullptr;  // FP32 master (kNParam)
  float* m = nullptr;
  float* v = nullptr;
  float* g = nullptr;
  uint64_t* w1p = nullptr;
  uint64_t* w2p = nullptr;
  uint64_t* w3p = nullptr;
  float* bias = nullptr;  // b1 (64), b2 (64), b3 (2)
  float* feat = nullptr;  // cap x 26
  float* targ = nullptr;  // cap x 2
  int32_t* count = nullptr;
  int32_t cap = 0;
  float* part = nullptr;
  float* lpart = nullptr;
  float* loss = nullptr;
  int32_t part_blocks = 0;
  int32_t adam_step = 0;
  hipStream_t stream = nullptr;
};

namespace {

MlpPacked packed(const bl_mlp* m) {
  MlpPacked W;
  W.w1 = m->w1p; W.w2 = m->w2p; W.w3 = m->w3p;
  W.b1 = m->bias; W.b2 = m->bias + kHid; W.b3 = m->bias + 2 * kHid;
  return W;
}

int mlp_pack(bl_mlp* m) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(4), dim3(256), 0, m->stream, m->w, m->w1p, m->w2p,
                     m->w3p, m->bias);
  return hipGetLastError() == hipSuccess ? BL_SUCCESS : BL_FAILURE;
}

int ensure_parts(bl_mlp* m, int nblk) {
  if (m->part_blocks >= nblk) return BL_SUCCESS;
  hipFree(m->part);
  hipFree(m->lpart);
  m->part = nullptr;
  m->lpart = nullptr;
  m->part_blocks = 0;
  if (hipMalloc(&m->part, sizeof(float) * (size_t)kNParam * nblk) != hipSuccess ||
      hipMalloc(&m->lpart, sizeof(float) * nblk) != hipSuccess) {
    set_error("mlp: out of device memory");
    return BL_FAILURE;
  }
  m->part_blocks = nblk;
  return BL_SUCCESS;
}

// gradient of the batch's loss into m->g, loss into m->loss
int mlp_gradient(bl_mlp* m, const int32_t* idx, int npairs, int batch, uint64_t seed, int it) {
  const int nblk = (batch + kTrainBlock - 1) / kTrainBlock;
  if (ensure_parts(m, nblk)) return BL_FAILURE;
  hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(kTrainBlock), 0, m->stream, m->w, m->feat,
                     m->targ, idx, npairs, batch, seed, it, m->part, m->lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, nblk,
                     m->part, m->lpart, m->g, m->loss, batch);
  return hipGetLastError() == hipSuccess ? BL_SUCCESS : BL_FAILURE;
}

}  // namespace

namespace bl {

int learned_velocity(Handle* hh, const double* x, const double* y, double* wu, double* wv,
                     double dt, const DevTimestep* ts) {
  bl_handle_t* h = static_cast<bl_handle_t*>(hh);
  const bl_mlp* m = h->mlp_infer;
  const int nnd = h->sz.nnd, nel = h->sz.nel, nr = h->nrelaxed;
  if (!m || !h->d_ring || !h->d_lpred) {
    set_error("learned relaxation: not set up (bl_relax_learned)");
    return BL_FAILURE;
  }
  const unsigned nbm = (unsigned)std::max((nr + 63) / 64, 1);
  BL_LAUNCH(h, "mlp_nodes", k_mlp_nodes, nbm, 256, packed(m), (const int32_t*)h->d_ring,
            (const int32_t*)h->d_rlist, nr, x, y, nnd, h->d_lpred, h->d_lh, h->d_lflag, ts);
  BL_LAUNCH(h, "learned_check", k_learned_check, nblocks(nel), kBlock, nel, nnd,
            (const int32_t*)h->I(BL_IELND), x, y, (const float*)h->d_lpred,
            (const double*)h->d_lh, h->d_lflag, ts);
  BL_LAUNCH(h, "learned_velocity", k_learned_velocity, nblocks(nnd), kBlock, nnd, x, y,
            (const float*)h->d_lpred, (const double*)h->d_lh, wu, wv, dt, h->d_lflag, ts);
  // the plain relaxation, whose kernels run only when the check failed
  return relax_velocity(h, x, y, wu, wv, dt, ts, h->d_lflag);
}

int collect_pairs(Handle* hh, const double* x, const double* y, const double* wu,
                  const double* wv, double dt, const DevTimestep* ts) {
  bl_handle_t* h = static_cast<bl_handle_t*>(hh);
  bl_mlp* m = h->mlp_collect;
  if (!m) return BL_SUCCESS;
  const int nr = h->nrelaxed, stride = h->collect_stride;
  const int take = (nr + stride - 1) / stride;
  BL_LAUNCH(h, "collect", k_collect, (unsigned)std::max((take + 255) / 256, 1), 256,
            (const int32_t*)h->d_ring, (const int32_t*)h->d_rlist, nr, stride, x, y, wu, wv, dt,
            (const int32_t*)m->count, m->cap, m->feat, m->targ, ts);
  BL_LAUNCH(h, "collect_end", k_collect_end, 1, 64, m->count, nr, stride, m->cap, ts);
  return BL_SUCCESS;
}

int learned_reset(Handle* h) {
  if (!h->d_lpred) return BL_SUCCESS;
  const size_t nnd = (size_t)std::max(h->sz.nnd, 1);
  BL_HIP(hipMemsetAsync(h->d_lpred, 0, sizeof(float) * 2 * nnd, h->stream));
  BL_HIP(hipMemsetAsync(h->d_lh, 0, sizeof(double) * nnd, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  return BL_SUCCESS;
}

int mesh_motion(Handle* h, const double* x, const double* y, double* wu, double* wv, double dt,
                const DevTimestep* ts) {
  if (h->mlp_infer) return learned_velocity(h, x, y, wu, wv, dt, ts);
  if (relax_velocity(h, x, y, wu, wv, dt, ts)) return BL_FAILURE;
  return collect_pairs(h, x, y, wu, wv, dt, ts);
}

}  // namespace bl

extern "C" {

int bl_mlp_create(uint64_t seed, int32_t capacity, bl_mlp_t** out) {
  if (!out || capacity < 0) { set_error("bl_mlp_create: bad argument"); return BL_FAILURE; }
  bl_mlp* m = new bl_mlp();
  m->cap = capacity;
  bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&m->w, sizeof(float) * kNParam) == hipSuccess &&
            hipMalloc(&m->m, sizeof(float) * kNParam) == hipSuccess &&
            hipMalloc(&m->v, sizeof(float) * kNParam) == hipSuccess &&
            hipMalloc(&m->g, sizeof(float) * kNParam) == hipSuccess &&
            hipMalloc(&m->w1p, sizeof(uint64_t) * 4 * 2 * 64) == hipSuccess &&
            hipMalloc(&m->w2p, sizeof(uint64_t) * 4 * 4 * 64) == hipSuccess &&
            hipMalloc(&m->w3p, sizeof(uint64_t) * 4 * 64) == hipSuccess &&
            hipMalloc(&m->bias, sizeof(float) * (2 * kHid + kOut)) == hipSuccess &&
            hipMalloc(&m->feat, sizeof(float) * kIn * (size_t)std::max(capacity, 1)) == hipSuccess &&
            hipMalloc(&m->targ, sizeof(float) * kOut * (size_t)std::max(capacity, 1)) == hipSuccess &&
            hipMalloc(&m->count, sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&m->loss, sizeof(float)) == hipSuccess;
  if (!ok) { set_error("bl_mlp_create: out of device memory"); bl_mlp_destroy(m); return BL_FAILURE; }
  // He-uniform initial weights from the seed (host), zero biases
  std::vector<float> w(kNParam, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int rows, int cols) {
    const double a = std::sqrt(6.0 / cols);
    std::uniform_real_distribution<double> u(-a, a);
    for (int i = 0; i < rows * cols; ++i) w[off + i] = (float)u(rng);
  };
  fill(kOffW1, kHid, kIn);
  fill(kOffW2, kHid, kHid);
  fill(kOffW3, kOut, kHid);
  std::vector<float> zero(kNParam, 0.0f);
  hipMemcpy(m->w, w.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice);
  hipMemcpy(m->m, zero.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice);
  hipMemcpy(m->v, zero.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice);
  hipMemset(m->count, 0, sizeof(int32_t));
  if (mlp_pack(m) || hipStreamSynchronize(m->stream) != hipSuccess) {
    set_error("bl_mlp_create: initialisation failed");
    bl_mlp_destroy(m);
    return BL_FAILURE;
  }
  *out = m;
  return BL_SUCCESS;
}

int bl_mlp_destroy(bl_mlp_t* m) {
  if (!m) return BL_SUCCESS;
  if (m->stream) hipStreamSynchronize(m->stream);
  void* bufs[] = {m->w, m->m, m->v, m->g, m->w1p, m->w2p, m->w3p, m->bias, m->feat, m->targ,
                  m->count, m->part, m->lpart, m->loss};
  for (void* p : bufs) if (p) hipFree(p);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return BL_SUCCESS;
}

int bl_mlp_weights(bl_mlp_t* m, float* get, const float* set) {
  if (!m) return BL_FAILURE;
  if (get) BL_HIP(hipMemcpy(get, m->w, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
  if (set) {
    BL_HIP(hipMemcpy(m->w, set, sizeof(float) * kNParam, hipMemcpyHostToDevice));
    if (mlp_pack(m)) return BL_FAILURE;
    BL_HIP(hipStreamSynchronize(m->stream));
  }
  return BL_SUCCESS;
}

int bl_mlp_forward(bl_mlp_t* m, const float* x, int32_t n, float* y) {
  if (!m || !x || !y || n < 0) return BL_FAILURE;
  if (n == 0) return BL_SUCCESS;
  float *dx = nullptr, *dy = nullptr;
  BL_HIP(hipMalloc(&dx, sizeof(float) * kIn * (size_t)n));
  BL_HIP(hipMalloc(&dy, sizeof(float) * kOut * (size_t)n));
  hipMemcpy(dx, x, sizeof(float) * kIn * (size_t)n, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + 63) / 64), dim3(256), 0, m->stream, packed(m),
                     (const float*)dx, n, dy);
  hipError_t e = hipGetLastError();
  hipStreamSynchronize(m->stream);
  hipMemcpy(y, dy, sizeof(float) * kOut * (size_t)n, hipMemcpyDeviceToHost);
  hipFree(dx);
  hipFree(dy);
  if (e != hipSuccess) { set_error("bl_mlp_forward: launch failed"); return BL_FAILURE; }
  return BL_SUCCESS;
}

int bl_mlp_add_pairs(bl_mlp_t* m, const float* x, const float* t, int32_t n) {
  if (!m || !x || !t || n < 0) return BL_FAILURE;
  int32_t c = 0;
  BL_HIP(hipMemcpy(&c, m->count, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (c + n > m->cap) { set_error("bl_mlp_add_pairs: capacity exceeded"); return BL_FAILURE; }
  BL_HIP(hipMemcpy(m->feat + (size_t)c * kIn, x, sizeof(float) * kIn * (size_t)n,
                   hipMemcpyHostToDevice));
  BL_HIP(hipMemcpy(m->targ + (size_t)c * kOut, t, sizeof(float) * kOut * (size_t)n,
                   hipMemcpyHostToDevice));
  c += n;
  BL_HIP(hipMemcpy(m->count, &c, sizeof(int32_t), hipMemcpyHostToDevice));
  return BL_SUCCESS;
}

int bl_mlp_pairs(bl_mlp_t* m, int32_t* n, float* x, float* t) {
  if (!m || !n) return BL_FAILURE;
  BL_HIP(hipStreamSynchronize(m->stream));
  BL_HIP(hipMemcpy(n, m->count, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (x) BL_HIP(hipMemcpy(x, m->feat, sizeof(float) * kIn * (size_t)*n, hipMemcpyDeviceToHost));
  if (t) BL_HIP(hipMemcpy(t, m->targ, sizeof(float) * kOut * (size_t)*n, hipMemcpyDeviceToHost));
  return BL_SUCCESS;
}

int bl_mlp_gradient(bl_mlp_t* m, int32_t n, float* grad, float* loss) {
  if (!m || n <= 0 || !grad || !loss) return BL_FAILURE;
  int32_t c = 0;
  BL_HIP(hipMemcpy(&c, m->count, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (n > c) { set_error("bl_mlp_gradient: more samples than pairs"); return BL_FAILURE; }
  std::vector<int32_t> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  int32_t* didx = nullptr;
  BL_HIP(hipMalloc(&didx, sizeof(int32_t) * n));
  hipMemcpy(didx, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
  int rc = mlp_gradient(m, didx, c, n, 0, 0);
  hipStreamSynchronize(m->stream);
  hipFree(didx);
  if (rc) return BL_FAILURE;
  BL_HIP(hipMemcpy(grad, m->g, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
  BL_HIP(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  return BL_SUCCESS;
}

int bl_mlp_train(bl_mlp_t* m, int32_t iters, int32_t batch, float lr, uint64_t seed,
                 float* loss_hist) {
  if (!m || iters < 0 || batch <= 0) return BL_FAILURE;
  int32_t c = 0;
  BL_HIP(hipStreamSynchronize(m->stream));
  BL_HIP(hipMemcpy(&c, m->count, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (c <= 0) { set_error("bl_mlp_train: no training pairs"); return BL_FAILURE; }
  std::vector<float> lh(iters);
  for (int it = 0; it < iters; ++it) {
    if (mlp_gradient(m, nullptr, c, batch, seed, it)) return BL_FAILURE;
    ++m->adam_step;
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->w,
                       m->m, m->v, (const float*)m->g, lr, m->adam_step);
    if (loss_hist)
      BL_HIP(hipMemcpyAsync(&lh[it], m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  }
  if (mlp_pack(m)) return BL_FAILURE;
  BL_HIP(hipStreamSynchronize(m->stream));
  if (loss_hist) std::copy(lh.begin(), lh.end(), loss_hist);
  return BL_SUCCESS;
}

int bl_relax_collect(bl_handle_t* h, bl_mlp_t* m, int32_t stride) {
  if (!h) return BL_FAILURE;
  if (m && (!h->d_ring || h->relax_type == BL_RELAX_OFF)) {
    set_error("bl_relax_collect: set a plain relaxation first (bl_set_relaxation)");
    return BL_FAILURE;
  }
  if (m && (stride < 1 || h->mlp_infer)) {
    set_error("bl_relax_collect: stride >= 1, and not while the learned relaxation is on");
    return BL_FAILURE;
  }
  BL_HIP(hipStreamSynchronize(h->stream));
  invalidate_graph(h);
  h->mlp_collect = m;
  h->collect_stride = m ? stride : 1;
  return BL_SUCCESS;
}

int bl_relax_learned(bl_handle_t* h, bl_mlp_t* m) {
  if (!h) return BL_FAILURE;
  if (m && (!h->d_ring || h->relax_type == BL_RELAX_OFF)) {
    set_error("bl_relax_learned: set the plain relaxation (the fallback) first");
    return BL_FAILURE;
  }
  if (m && h->mlp_collect) {
    set_error("bl_relax_learned: not while pairs are being collected (bl_relax_collect)");
    return BL_FAILURE;
  }
  BL_HIP(hipStreamSynchronize(h->stream));
  invalidate_graph(h);
  h->mlp_infer = m;
  if (m) {
    if (!h->d_lpred) BL_HIP(hipMalloc(&h->d_lpred, sizeof(float) * 2 * std::max(h->sz.nnd, 1)));
    if (!h->d_lh) BL_HIP(hipMalloc(&h->d_lh, sizeof(double) * std::max(h->sz.nnd, 1)));
    if (!h->d_lflag) {
      BL_HIP(hipMalloc(&h->d_lflag, sizeof(int32_t) * 4));
      BL_HIP(hipMemset(h->d_lflag, 0, sizeof(int32_t) * 4));
    }
    if (learned_reset(h)) return BL_FAILURE;
  }
  return BL_SUCCESS;
}

int bl_relax_learned_stats(bl_handle_t* h, int32_t* steps, int32_t* fallbacks) {
  if (!h) return BL_FAILURE;
  int32_t f[4] = {0, 0, 0, 0};
  if (h->d_lflag) {
    BL_HIP(hipStreamSynchronize(h->stream));
    BL_HIP(hipMemcpy(f, h->d_lflag, sizeof(int32_t) * 4, hipMemcpyDeviceToHost));
  }
  if (steps) *steps = f[1];
  if (fallbacks) *fallbacks = f[2];
  return BL_SUCCESS;
}

// the features of every relaxed node at the current positions (nnd x 26,
// zero for nodes that are not relaxed)
int bl_relax_features(bl_handle_t* h, float* feat) {
  if (!h || !feat || !h->d_ring || !h->d_rlist) {
    set_error("bl_relax_features: set a relaxation first");
    return BL_FAILURE;
  }
  const int nnd = h->sz.nnd, nr = h->nrelaxed;
  float* df = nullptr;
  float* dt = nullptr;
  int32_t* dc = nullptr;
  double* zero = nullptr;
  BL_HIP(hipStreamSynchronize(h->stream));
  BL_HIP(hipMalloc(&df, sizeof(float) * kIn * std::max(nr, 1)));
  BL_HIP(hipMalloc(&dt, sizeof(float) * kOut * std::max(nr, 1)));
  BL_HIP(hipMalloc(&dc, sizeof(int32_t)));
  BL_HIP(hipMalloc(&zero, sizeof(double) * std::max(nnd, 1)));
  BL_HIP(hipMemset(dc, 0, sizeof(int32_t)));
  BL_HIP(hipMemset(zero, 0, sizeof(double) * std::max(nnd, 1)));
  // the collect kernel with zero velocity and stride 1 gives the features
  hipLaunchKernelGGL(k_collect, dim3(std::max((nr + 255) / 256, 1)), dim3(256), 0, h->stream,
                     (const int32_t*)h->d_ring, (const int32_t*)h->d_rlist, nr, 1,
                     (const double*)h->R(BL_NDX), (const double*)h->R(BL_NDY),
                     (const double*)zero, (const double*)zero, 1.0, (const int32_t*)dc, nr, df,
                     dt, (const DevTimestep*)nullptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  std::vector<float> f((size_t)kIn * std::max(nr, 1));
  std::vector<int32_t> rl(std::max(nr, 1));
  if (e == hipSuccess) e = hipMemcpy(f.data(), df, sizeof(float) * kIn * nr, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(rl.data(), h->d_rlist, sizeof(int32_t) * nr, hipMemcpyDeviceToHost);
  hipFree(df);
  hipFree(dt);
  hipFree(dc);
  hipFree(zero);
  if (e != hipSuccess) {
    set_error(std::string("bl_relax_features: ") + hipGetErrorString(e));
    return BL_FAILURE;
  }
  std::fill(feat, feat + (size_t)kIn * nnd, 0.0f);
  for (int i = 0; i < nr; ++i)
    std::copy(&f[(size_t)i * kIn], &f[(size_t)i * kIn] + kIn, feat + (size_t)rl[i] * kIn);
  return BL_SUCCESS;
}

}  // extern "C"
