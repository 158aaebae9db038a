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
//   f[24..25] the displacement of one sweep of the relaxation / h (the
//             stencil weights of the relaxation set on the handle)
// Target: the plain relaxation's displacement (x_R - x, y_R - y)/h after its
// nsweep sweeps; the network learns the difference from f[24..25], so that
// a network that outputs zero is one sweep of the plain relaxation.
//
// Inference is one kernel: each lane computes one node's features, and
// each wave runs its 64 nodes through the three layers on the matrix cores
// (v_mfma_f32_16x16x16_bf16: BF16 weights and activations, FP32
// accumulation), 16 nodes per pass, weights and activations in registers.
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
#include "bl_relax.h"

namespace bl {

constexpr int kIn = 26, kInPad = 32, kHid = 64, kOut = 2;
// parameter layout (FP32 master): W1 (64,26) row-major, b1, W2 (64,64), b2,
// W3 (2,64), b3
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid,
              kOffB2 = kOffW2 + kHid * kHid, kOffW3 = kOffB2 + kHid,
              kOffB3 = kOffW3 + kOut * kHid, kNParam = kOffB3 + kOut;
constexpr int kTrainBlock = 64;  // samples per training workgroup
// learned-motion flags: [0] this step falls back, [1] steps, [2] fallbacks,
// [3] reasons seen, [4] this step used the guard, [5] guarded steps
constexpr int kLFlags = 8;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// round to nearest even BF16 (the hardware conversion) and back
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t b) {
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}
__device__ __forceinline__ float bfr(float f) { return bf2f(f2bf(f)); }
// two floats to two BF16 in one 32-bit word (the first in the low half)
__device__ __forceinline__ uint32_t pk2bf(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{a, b}, bf16x2_t));
}

// The features of a relaxed node: the neighbour offsets in FP64, the rest
// in FP32 on the offsets (the network sees them rounded to BF16); h the
// length scale.  The one-sweep displacement [24..25] uses the relaxation's
// stencil weights (setup_mesh.F90:804-857 for EQUI; LIN1/LIN2 the two
// l- or k-neighbours) on the normalised offsets: om * sum r_i d_i / h, the
// node staying put when neither residual exceeds tol or, for EQUI, when
// |zeta| < zerocut, as in one plain sweep.
__device__ __forceinline__ void node_features(const RelaxArgs& R, const double* __restrict__ x,
                                              const double* __restrict__ y, int n,
                                              float f[kIn], float& h) {
  const int4 qa = reinterpret_cast<const int4*>(R.ring)[2 * n];
  const int4 qb = reinterpret_cast<const int4*>(R.ring)[2 * n + 1];
  const int nb[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
  const double xn = x[n], yn = y[n];
  float dx[8], dy[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    dx[i] = (float)(x[nb[i]] - xn);
    dy[i] = (float)(y[nb[i]] - yn);
  }
  // counter-clockwise ring P0..P7 (P0 = +k) from the table order
  const int ccw[8] = {4, 2, 1, 0, 3, 5, 6, 7};
  float area[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int a = ccw[2 * q], b = ccw[2 * q + 1], c = ccw[(2 * q + 2) & 7];
    area[q] = 0.5f * (dx[b] * (dy[c] - dy[a]) - (dx[c] - dx[a]) * dy[b]);
  }
  const float h2 = (area[0] + area[1] + area[2] + area[3]) / 4.0f;
  h = sqrtf(h2 > 1.0e-30f ? h2 : 1.0e-30f);
  const float ih = 1.0f / h;
  float u[8], v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    u[i] = dx[i] * ih;
    v[i] = dy[i] * ih;
    f[2 * i] = u[i];
    f[2 * i + 1] = v[i];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int a = ccw[2 * q], b = ccw[2 * q + 1], c = ccw[(2 * q + 2) & 7];
    f[16 + q] = 0.5f * (u[b] * (v[c] - v[a]) - (u[c] - u[a]) * v[b]);
    f[20 + q] = u[a] * v[c] - u[c] * v[a];
  }
  // one sweep: weights r_i in table order, normalised to sum 1
  float r[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  bool moves = true;
  if (R.type == BL_RELAX_EQUI) {
    const float r_psi = 0.5f * (v[4] - v[3]), s_psi = 0.5f * (u[4] - u[3]);
    const float r_phi = 0.5f * (v[1] - v[6]), s_phi = 0.5f * (u[1] - u[6]);
    const float zeta = s_psi * r_phi - s_phi * r_psi;
    const float alph = s_psi * s_psi + r_psi * r_psi;
    const float beta = s_phi * s_psi + r_phi * r_psi;
    const float gamm = s_phi * s_phi + r_phi * r_phi;
    moves = !((double)fabsf(zeta) * ((double)h * (double)h) < R.zerocut);
    const float fac = 2.0f * (alph + gamm);
    r[0] = 0.5f * beta / fac; r[1] = alph / fac; r[2] = -r[0]; r[3] = gamm / fac;
    r[4] = r[3]; r[5] = r[2]; r[6] = r[1]; r[7] = r[0];
  } else if (R.type == BL_RELAX_LIN1) {
    r[1] = 0.5f; r[6] = 0.5f;
  } else {
    r[3] = 0.5f; r[4] = 0.5f;
  }
  float su = 0.0f, sv = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    su = su + r[i] * u[i];
    sv = sv + r[i] * v[i];
  }
  moves = moves && ((double)fabsf(sv) * h > R.tol || (double)fabsf(su) * h > R.tol);
  f[24] = moves ? (float)R.om * su : 0.0f;
  f[25] = moves ? (float)R.om * sv : 0.0f;
}

// The network runs transposed on the matrix cores: for 16 nodes at a time,
// Y^T (outputs x nodes) = W (outputs x inputs) X^T (inputs x nodes), with
// v_mfma_f32_16x16x32_bf16 taking a 16x32 weight tile as its A operand and
// a 32-input slice of the activations as its B operand.  Operand layouts
// (wave64): A lane l holds A[l%16][8(l/16)+j], B lane l holds
// B[8(l/16)+j][l%16] (j = 0..7), the result lane l holds D[4(l/16)+j][l%16]
// (j = 0..3).  The k index of a slice is free as long as A and B agree: we
// take hardware k = 8g + j to mean input 4g + j (j < 4) or 16 + 4g + j - 4
// (j >= 4) of the slice, so that lane l's results of output tiles 2s and
// 2s+1, side by side, are exactly its B operand of the next layer's input
// slice s.  The activations stay in registers from layer to layer.
//
// Packed weights: one 16-byte group of eight BF16 per (tile t, slice s,
// lane l), element j = W[16t + l%16][32s + perm(8(l/16) + j)] (zero past
// the layer's size).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

constexpr int kW1Groups = 4, kW2Groups = 8, kW3Groups = 2;  // (tile, slice) groups
__host__ __device__ constexpr int slice_input(int k) {     // hardware k -> input
  return (k & 7) < 4 ? 4 * (k >> 3) + (k & 7) : 16 + 4 * (k >> 3) + (k & 7) - 4;
}

struct MlpPacked {
  const u64x2* w1;  // (4 tiles, 64 lanes)
  const u64x2* w2;  // (4 tiles x 2 slices, 64)
  const u64x2* w3;  // (2 slices, 64), rows 0,1 real
  const float* b1;
  const float* b2;
  const float* b3;
};

// the network staged in LDS by each workgroup: lane-ordered A operands and
// the biases as accumulator initial values, per (tile, lane group)
struct MlpShared {
  u64x2 w1[kW1Groups][64], w2[kW2Groups][64], w3[kW3Groups][64];
  f32x4 b1[4][4], b2[4][4], b3[4];
};

__device__ __forceinline__ void stage_network(const MlpPacked& W, MlpShared& S) {
  const int t = threadIdx.x;
  for (int i = t; i < kW1Groups * 64; i += blockDim.x) S.w1[i / 64][i % 64] = W.w1[i];
  for (int i = t; i < kW2Groups * 64; i += blockDim.x) S.w2[i / 64][i % 64] = W.w2[i];
  for (int i = t; i < kW3Groups * 64; i += blockDim.x) S.w3[i / 64][i % 64] = W.w3[i];
  if (t < 64) {
    const int tile = t / 16, g = (t / 4) % 4, j = t % 4;
    S.b1[tile][g][j] = W.b1[16 * tile + 4 * g + j];
    S.b2[tile][g][j] = W.b2[16 * tile + 4 * g + j];
  }
  if (t < 16) S.b3[t / 4][t % 4] = t < kOut ? W.b3[t] : 0.0f;
}

__device__ __forceinline__ bf16x8 cat_bf16(uint64_t lo, uint64_t hi) {
  return __builtin_bit_cast(bf16x8, u64x2{lo, hi});
}
__device__ __forceinline__ uint64_t relu_pack4(const f32x4 z) {
  const uint32_t lo = pk2bf(fmaxf(z[0], 0.0f), fmaxf(z[1], 0.0f));
  const uint32_t hi = pk2bf(fmaxf(z[2], 0.0f), fmaxf(z[3], 0.0f));
  return (uint64_t)lo | ((uint64_t)hi << 32);
}
__device__ __forceinline__ f32x4 mfma32(const u64x2 a, const bf16x8 b, const f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), b, c, 0, 0, 0);
}

// the three layers for the wave's 64 nodes, as four groups of 16 taken
// layer by layer (each weight fragment read from LDS once per wave, four
// independent accumulations per fragment): in[q] the B operand of group
// q's 32 (padded) inputs; lane l (l < 16) gets node 16q + l's two outputs
// in y[q][0], y[q][1]
__device__ __forceinline__ void mlp_wave(const MlpShared& S, const bf16x8 in[4], f32x4 y[4]) {
  const int l = threadIdx.x & 63, g = l >> 4;
  uint64_t a1[4][4], a2[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const u64x2 w = S.w1[t][l];
    const f32x4 b = S.b1[t][g];
#pragma unroll
    for (int q = 0; q < 4; ++q) a1[q][t] = relu_pack4(mfma32(w, in[q], b));
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const u64x2 w0 = S.w2[2 * t][l], w1 = S.w2[2 * t + 1][l];
    const f32x4 b = S.b2[t][g];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const f32x4 acc = mfma32(w0, cat_bf16(a1[q][0], a1[q][1]), b);
      a2[q][t] = relu_pack4(mfma32(w1, cat_bf16(a1[q][2], a1[q][3]), acc));
    }
  }
  const u64x2 w0 = S.w3[0][l], w1 = S.w3[1][l];
  const f32x4 b = S.b3[g];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    y[q] = mfma32(w1, cat_bf16(a2[q][2], a2[q][3]), mfma32(w0, cat_bf16(a2[q][0], a2[q][1]), b));
}

// The standardisation of the network's inputs and outputs (FP32, applied
// before the BF16 rounding so that small deviations of the stencil from its
// mean survive): norm[0..25] input means, norm[26..51] inverse input
// standard deviations, norm[52..53] output means, norm[54..55] output
// standard deviations.  The network maps (f - mu) * isig to (t - mu_t)/sig_t.
constexpr int kNorm = 2 * kIn + 2 * kOut;
constexpr int kNormMu = 0, kNormIsig = kIn, kNormMuT = 2 * kIn, kNormSigT = 2 * kIn + kOut;

__device__ __forceinline__ float norm_in(const float* __restrict__ norm, int k, float f) {
  return (f - norm[kNormMu + k]) * norm[kNormIsig + k];
}
__device__ __forceinline__ float norm_out(const float* __restrict__ norm, int c, float y) {
  return norm[kNormMuT + c] + norm[kNormSigT + c] * y;
}

// 64 rows per wave: lane l's row f (26 features, standardised and rounded
// to BF16 here) goes through the network; returns lane l's raw outputs.
// xw (64 x 8 groups of four BF16) and ys (64) are the wave's LDS rows; every
// wave of the workgroup calls it (workgroup barriers, which also cover the
// staging of S).
__device__ __forceinline__ float2 mlp_rows64(const MlpShared& S, const float* __restrict__ norm,
                                             const float f[kIn], uint64_t* xw, float2* ys) {
  const int l = threadIdx.x & 63, col = l & 15, grp = l >> 4;
#pragma unroll
  for (int g = 0; g < kInPad / 4; ++g) {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = 4 * g + j < kIn ? norm_in(norm, 4 * g + j, f[4 * g + j]) : 0.0f;
    xw[l * (kInPad / 4) + g] = (uint64_t)pk2bf(v[0], v[1]) | ((uint64_t)pk2bf(v[2], v[3]) << 32);
  }
  __syncthreads();
  bf16x8 in[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // inputs 4grp..4grp+3 and 16+4grp..16+4grp+3 of node 16q + col
    const uint64_t* row = xw + (16 * q + col) * (kInPad / 4);
    in[q] = cat_bf16(row[grp], row[4 + grp]);
  }
  f32x4 y[4];
  mlp_wave(S, in, y);
  if (grp == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) ys[16 * q + col] = make_float2(y[q][0], y[q][1]);
  }
  __syncthreads();
  return ys[l];
}

// inference on a feature buffer (n x 26 FP32) into y (n x 2), the
// displacement / h f[24..25] + the network's difference: 256 rows per
// workgroup
__global__ void __launch_bounds__(256) k_mlp_forward(MlpPacked W, const float* __restrict__ norm,
                                                     const float* __restrict__ xin, int n,
                                                     float* __restrict__ yout) {
  __shared__ MlpShared S;
  __shared__ uint64_t xs[4][64 * (kInPad / 4)];
  __shared__ float2 ys[4][64];
  const int wave = threadIdx.x >> 6;
  const int r = blockIdx.x * 256 + threadIdx.x;
  stage_network(W, S);
  float f[kIn];
#pragma unroll
  for (int k = 0; k < kIn; ++k) f[k] = r < n ? xin[(size_t)r * kIn + k] : 0.0f;
  const float2 y = mlp_rows64(S, norm, f, xs[wave], ys[wave]);
  if (r < n) {
    yout[(size_t)r * kOut] = f[24] + norm_out(norm, 0, y.x);
    yout[(size_t)r * kOut + 1] = f[25] + norm_out(norm, 1, y.y);
  }
}

// The learned mesh velocity of a step: one relaxed node per lane (features
// in FP32 on FP64 offsets, standardised), the relaxed nodes in nchunk chunks
// of 256, workgroup b < nbm taking chunks b, b + nbm, ... with the network
// staged in its LDS once and run on the matrix cores, w = h t / dt with t =
// f[24..25] + the network's difference, the displacement / h; the one-sweep
// displacement h f[24..25] into d1 (the check's reference and the guard's
// motion) and the node's guard mark cleared.  Workgroups from nbm set w and
// d1 to zero on the nodes that are not relaxed.  Workgroup 0 also keeps the
// step counters: the previous step's fallback flag[0] and guard flag[4] are
// added to their counts flag[2] and flag[5] and cleared, flag[1] counts
// steps.
constexpr int kMlpBlocks = 1024;  // four per CU, all resident at once

__global__ void __launch_bounds__(256) k_mlp_nodes(MlpPacked W, const float* __restrict__ norm,
                                                   RelaxArgs RA,
                                                   const int32_t* __restrict__ rlist, int nr,
                                                   const int32_t* __restrict__ zlist, int nz,
                                                   int nchunk, int nbm,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   double* __restrict__ wu,
                                                   double* __restrict__ wv, double dt,
                                                   float2* __restrict__ d1,
                                                   int32_t* __restrict__ mark,
                                                   int32_t* __restrict__ flag,
                                                   const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  if (ts) dt = ts->dt;
  if ((int)blockIdx.x >= nbm) {
    const int i = ((int)blockIdx.x - nbm) * blockDim.x + threadIdx.x;
    if (i < nz) {
      const int n = zlist[i];
      wu[n] = 0.0;
      wv[n] = 0.0;
      d1[n] = make_float2(0.0f, 0.0f);
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (flag[0]) flag[2] = flag[2] + 1;
    if (flag[4]) flag[5] = flag[5] + 1;
    flag[0] = 0;
    flag[4] = 0;
    flag[1] = flag[1] + 1;
  }
  __shared__ MlpShared S;
  __shared__ uint64_t xs[4][64 * (kInPad / 4)];
  __shared__ float2 ys[4][64];
  const int wave = threadIdx.x >> 6;
  // staged once, for every chunk of 256 relaxed nodes this workgroup takes
  stage_network(W, S);
  const double idt = 1.0 / dt;
  for (int chunk = blockIdx.x; chunk < nchunk; chunk += nbm) {
    const int r = chunk * 256 + threadIdx.x;
    const int node = r < nr ? rlist[r] : -1;
    float f[kIn];
    float h = 0.0f;
    if (node >= 0) {
      node_features(RA, x, y, node, f, h);
    } else {
#pragma unroll
      for (int k = 0; k < kIn; ++k) f[k] = 0.0f;
    }
    const float2 t = mlp_rows64(S, norm, f, xs[wave], ys[wave]);
    if (node >= 0) {
      mark[node] = 0;
      d1[node] = make_float2(h * f[24], h * f[25]);
      wu[node] = (double)h * (double)(f[24] + norm_out(norm, 0, t.x)) * idt;
      wv[node] = (double)h * (double)(f[25] + norm_out(norm, 1, t.y)) * idt;
    }
  }
}

// smallest corner Jacobian, squared longest/shortest side ratio and area
// of the quadrilateral (px, py), corners counter-clockwise
__device__ __forceinline__ void quad_quality(const double px[4], const double py[4], double& jmin,
                                             double& asp2, double& area) {
  double lmin = DBL_MAX, lmax = 0.0;
  jmin = DBL_MAX;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int a = (c + 3) & 3, b = (c + 1) & 3;
    const double j = (px[b] - px[c]) * (py[a] - py[c]) - (px[a] - px[c]) * (py[b] - py[c]);
    jmin = j < jmin ? j : jmin;
    const double ex = px[b] - px[c], ey = py[b] - py[c];
    const double len2 = ex * ex + ey * ey;
    lmin = len2 < lmin ? len2 : lmin;
    lmax = len2 > lmax ? len2 : lmax;
  }
  asp2 = lmin > 0.0 ? lmax / lmin : DBL_MAX;
  area = 0.5 * ((px[2] - px[0]) * (py[3] - py[1]) - (px[3] - px[1]) * (py[2] - py[0]));
}

// one element of the check below
__device__ __forceinline__ void check_element(int e, const int32_t* __restrict__ ielnd,
                                              const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ wu,
                                              const double* __restrict__ wv,
                                              const float2* __restrict__ d1, double dt,
                                              double amax2, int relative, int stage,
                                              int32_t* __restrict__ mark,
                                              int32_t* __restrict__ flag,
                                              bool guarded = false) {
  const int4 q = reinterpret_cast<const int4*>(ielnd)[e];
  const int nd[4] = {q.x - 1, q.y - 1, q.z - 1, q.w - 1};
  double px[4], py[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    // with the guard folded in, a marked node moves by d1/dt, the value the
    // guard writes (a node off the relaxed list has d1 = 0 and w = 0: the
    // same bits either way); its w is not read, as the guard may be
    // writing it
    double u, v;
    if (guarded && mark[nd[c]]) {
      const float2 d = d1[nd[c]];
      u = (double)d.x / dt;
      v = (double)d.y / dt;
    } else {
      u = wu[nd[c]];
      v = wv[nd[c]];
    }
    px[c] = x[nd[c]] + u * dt;
    py[c] = y[nd[c]] + v * dt;
  }
  double jp, ap, vp;
  quad_quality(px, py, jp, ap, vp);
  bool jbad = !(jp > 0.0), abad = !(ap <= amax2), vbad = false;
  if (relative) {
    double ox[4], oy[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float2 d = d1[nd[c]];
      ox[c] = x[nd[c]] + ((double)d.x / dt) * dt;
      oy[c] = y[nd[c]] + ((double)d.y / dt) * dt;
    }
    double jo, ao, vo;
    quad_quality(ox, oy, jo, ao, vo);
    jbad = jbad && jp < jo;
    abad = abad && ap > ao;
    // an element the proposal shrinks below half its one-sweep area
    vbad = !(vp >= 0.5 * vo);
  }
  const int code = (jbad ? 1 : 0) | (abad ? 2 : 0) | (vbad ? 4 : 0);
  if (!code) return;
  atomicOr(&flag[3], code);
  if (stage == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) mark[nd[c]] = 1;
    flag[4] = 1;
  } else {
    atomicOr(&flag[0], code);
  }
}

// The quality check of the proposed mesh x + w dt: an element fails with a
// corner Jacobian <= 0 (bit 0) or a longest/shortest side ratio above
// sqrt(amax2) (bit 1).  With relative set, such an element counts only when
// the proposal is worse than the same element after one plain sweep
// x + (d1/dt) dt (what the guard would give it), and an element also fails
// when the proposal shrinks it below half its one-sweep area (bit 2).  stage 0: a counted element
// marks its four nodes for the guard and sets flag[4]; stage 1 (only when
// flag[4] is set, with the guard's velocities: k_learned_recheck): a counted
// element sets the step's fallback flag[0].  flag[3] keeps every reason seen
// since the network was attached.
__global__ void k_learned_check(int nel, const int32_t* __restrict__ ielnd,
                                const double* __restrict__ x, const double* __restrict__ y,
                                const double* __restrict__ wu, const double* __restrict__ wv,
                                const float2* __restrict__ d1, double dt, double amax2,
                                int relative, int stage, int32_t* __restrict__ mark,
                                int32_t* __restrict__ flag, const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  if (ts) dt = ts->dt;
  const int stride = gridDim.x * blockDim.x;
  for (int e = xcd_index(); e < nel; e += stride)
    check_element(e, ielnd, x, y, wu, wv, d1, dt, amax2, relative, stage, mark, flag);
}

// Stage 1 with the guard folded in (only when stage 0 marked nodes,
// flag[4]): a relaxed node marked by stage 0 moves by one plain sweep
// instead of the network's proposal (written here), and every element is
// checked again with the guarded velocities, which it forms itself from the
// marks, so that the check needs no pass of its own after the guard's
__global__ void k_learned_recheck(int nel, const int32_t* __restrict__ ielnd,
                                  const double* __restrict__ x, const double* __restrict__ y,
                                  double* __restrict__ wu, double* __restrict__ wv,
                                  const float2* __restrict__ d1, double dt, double amax2,
                                  int relative, const int32_t* __restrict__ rlist, int nr,
                                  int32_t* __restrict__ mark, int32_t* __restrict__ flag,
                                  const DevTimestep* ts) {
  if (ts && (ts->skip || !ts->ale_on)) return;
  if (flag[4] == 0) return;
  if (ts) dt = ts->dt;
  const int stride = gridDim.x * blockDim.x;
  for (int i = xcd_index(); i < nr; i += stride) {
    const int n = rlist[i];
    if (!mark[n]) continue;
    const float2 d = d1[n];
    wu[n] = (double)d.x / dt;
    wv[n] = (double)d.y / dt;
  }
  for (int e = xcd_index(); e < nel; e += stride)
    check_element(e, ielnd, x, y, wu, wv, d1, dt, amax2, relative, 1, mark, flag, true);
}

// per-column mean and standard deviation of the stored pairs (FP64 sums in
// a fixed order): one workgroup per column, 26 feature columns then the 2
// columns of target - f[24..25]
constexpr int kStatBlock = 256;

__global__ void __launch_bounds__(kStatBlock)
k_mlp_stats(const float* __restrict__ feat, const float* __restrict__ targ, int n,
            float* __restrict__ norm) {
  __shared__ double s1[kStatBlock], s2[kStatBlock];
  const int c = blockIdx.x, t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int i = t; i < n; i += kStatBlock) {
    // targets: their difference from one sweep, as the network learns it
    const double v = c < kIn ? (double)feat[(size_t)i * kIn + c]
                             : (double)(targ[(size_t)i * kOut + (c - kIn)] -
                                        feat[(size_t)i * kIn + 24 + (c - kIn)]);
    a = a + v;
    b = b + v * v;
  }
  s1[t] = a;
  s2[t] = b;
  __syncthreads();
  for (int w = kStatBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      s1[t] = s1[t] + s1[t + w];
      s2[t] = s2[t] + s2[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double mean = n > 0 ? s1[0] / n : 0.0;
    const double var = n > 0 ? s2[0] / n - mean * mean : 0.0;
    const double sd = var > 0.0 ? sqrt(var) : 0.0;
    if (c < kIn) {
      norm[kNormMu + c] = (float)mean;
      norm[kNormIsig + c] = sd > 1.0e-8 ? (float)(1.0 / sd) : 1.0f;
    } else {
      norm[kNormMuT + c - kIn] = (float)mean;
      norm[kNormSigT + c - kIn] = sd > 1.0e-12 ? (float)sd : 1.0f;
    }
  }
}

// training pairs from the plain relaxation's mesh velocity: features at x,
// target w * dt / h; every stride-th relaxed node, starting at the step
// number modulo stride, appended at pairs[count] while they fit
__device__ __forceinline__ int collect_phase(int stride, const DevTimestep* ts) {
  return ts ? ts->nstep % stride : 0;
}

__global__ void k_collect(RelaxArgs RA, const int32_t* __restrict__ rlist, int nr, int stride, const double* __restrict__ x,
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
  float h;
  node_features(RA, x, y, n, f, h);
  float* fo = feat + (size_t)(base + i) * kIn;
#pragma unroll
  for (int k = 0; k < kIn; ++k) fo[k] = f[k];
  targ[(size_t)(base + i) * kOut] = (float)(wu[n] * dt / (double)h);
  targ[(size_t)(base + i) * kOut + 1] = (float)(wv[n] * dt / (double)h);
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
k_mlp_train(const float* __restrict__ w, const float* __restrict__ norm,
            const float* __restrict__ feat,
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
  for (int k = 0; k < kIn; ++k)
    sx[t][k] = live ? bfr(norm_in(norm, k, feat[(size_t)src * kIn + k])) : 0.0f;
  // the difference of the target from one sweep f[24..25], standardised
  const float t0 =
      live ? ((targ[(size_t)src * kOut] - feat[(size_t)src * kIn + 24]) - norm[kNormMuT]) /
                 norm[kNormSigT]
           : 0.0f;
  const float t1 =
      live ? ((targ[(size_t)src * kOut + 1] - feat[(size_t)src * kIn + 25]) - norm[kNormMuT + 1]) /
                 norm[kNormSigT + 1]
           : 0.0f;
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

// pack the master weights as BF16 A operands (the layout of MlpPacked)
// and FP32 biases
__global__ void k_mlp_pack(const float* __restrict__ w, u64x2* __restrict__ w1p,
                           u64x2* __restrict__ w2p, u64x2* __restrict__ w3p,
                           float* __restrict__ bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  auto pack = [&](int rows, int cols, int off, int t, int sl, int l) {
    const int row = 16 * t + (l & 15);
    uint64_t v[2] = {0, 0};
    for (int j = 0; j < 8; ++j) {
      const int c = 32 * sl + slice_input(8 * (l >> 4) + j);
      const float f = (row < rows && c < cols) ? w[off + row * cols + c] : 0.0f;
      v[j >> 2] |= (uint64_t)f2bf(f) << (16 * (j & 3));
    }
    return u64x2{v[0], v[1]};
  };
  if (i < kW1Groups * 64) w1p[i] = pack(kHid, kIn, kOffW1, i / 64, 0, i % 64);
  if (i < kW2Groups * 64) w2p[i] = pack(kHid, kHid, kOffW2, i / 128, (i / 64) % 2, i % 64);
  if (i < kW3Groups * 64) w3p[i] = pack(kOut, kHid, kOffW3, 0, i / 64, i % 64);
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
  u64x2* w1p = nullptr;
  u64x2* w2p = nullptr;
  u64x2* w3p = nullptr;
  float* bias = nullptr;  // b1 (64), b2 (64), b3 (2)
  float* norm = nullptr;  // kNorm: input and output standardisation
  bool norm_set = false;  // set from pairs or by the caller
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

int mlp_normalize(bl_mlp* m) {
  int32_t c = 0;
  BL_HIP(hipStreamSynchronize(m->stream));
  BL_HIP(hipMemcpy(&c, m->count, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (c <= 0) {
    set_error("mlp: no pairs to standardise from");
    return BL_FAILURE;
  }
  hipLaunchKernelGGL(k_mlp_stats, dim3(kIn + kOut), dim3(kStatBlock), 0, m->stream,
                     (const float*)m->feat, (const float*)m->targ, c, m->norm);
  BL_HIP(hipGetLastError());
  BL_HIP(hipStreamSynchronize(m->stream));
  m->norm_set = true;
  return BL_SUCCESS;
}

int mlp_pack(bl_mlp* m) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(2), dim3(256), 0, m->stream, m->w, m->w1p, m->w2p,
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
  hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(kTrainBlock), 0, m->stream, m->w,
                     (const float*)m->norm, m->feat, m->targ, idx, npairs, batch, seed, it,
                     m->part, m->lpart);
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
  const int nel = h->sz.nel, nr = h->nrelaxed, nz = h->nzero;
  if (!m || !h->d_ring || !h->d_rlist || !h->d_zlist || !h->d_lflag || !h->d_lmark ||
      !h->d_ld1) {
    set_error("learned relaxation: not set up (bl_relax_learned)");
    return BL_FAILURE;
  }
  // at most kMlpBlocks workgroups, each taking every kMlpBlocks-th chunk of
  // 256 relaxed nodes (the network is staged once per workgroup), then the
  // zero workgroups
  const int nchunk = std::max((nr + 255) / 256, 1);
  const int nbm = std::min(nchunk, kMlpBlocks);
  const int nbz = (nz + 255) / 256;
  BL_LAUNCH(h, "mlp_nodes", k_mlp_nodes, (unsigned)(nbm + nbz), 256, packed(m),
            (const float*)m->norm, relax_args(h), (const int32_t*)h->d_rlist, nr,
            (const int32_t*)h->d_zlist, nz, nchunk, nbm, x, y, wu, wv, dt, (float2*)h->d_ld1,
            h->d_lmark, h->d_lflag, ts);
  const double amax2 = h->learned_max_aspect * h->learned_max_aspect;
  BL_LAUNCH(h, "learned_check", k_learned_check, nblocks(nel), kBlock, nel,
            (const int32_t*)h->I(BL_IELND), x, y, (const double*)wu, (const double*)wv,
            (const float2*)h->d_ld1, dt, amax2, h->learned_relative, 0, h->d_lmark,
            h->d_lflag, ts);
  BL_LAUNCH(h, "learned_recheck", k_learned_recheck,
            std::min(nblocks(std::max(nel, nr)), kGatedBlocks), kBlock, nel,
            (const int32_t*)h->I(BL_IELND), x, y, wu, wv, (const float2*)h->d_ld1, dt, amax2,
            h->learned_relative, (const int32_t*)h->d_rlist, nr, h->d_lmark, h->d_lflag, ts);
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
            relax_args(h), (const int32_t*)h->d_rlist, nr, stride, x, y, wu, wv, dt,
            (const int32_t*)m->count, m->cap, m->feat, m->targ, ts);
  BL_LAUNCH(h, "collect_end", k_collect_end, 1, 64, m->count, nr, stride, m->cap, ts);
  return BL_SUCCESS;
}

int mesh_motion(Handle* h, const double* x, const double* y, double* wu, double* wv, double dt,
                const DevTimestep* ts) {
  if (h->mlp_infer) {
    // collecting while the network moves the mesh: the plain relaxation's
    // motion of the same Lagrangian mesh, into buffers of its own, is the
    // target (states the network itself leads to)
    if (h->mlp_collect) {
      if (relax_velocity(h, x, y, h->d_cw[0], h->d_cw[1], dt, ts) ||
          collect_pairs(h, x, y, h->d_cw[0], h->d_cw[1], dt, ts))
        return BL_FAILURE;
    }
    return learned_velocity(h, x, y, wu, wv, dt, ts);
  }
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
            hipMalloc(&m->w1p, sizeof(u64x2) * kW1Groups * 64) == hipSuccess &&
            hipMalloc(&m->w2p, sizeof(u64x2) * kW2Groups * 64) == hipSuccess &&
            hipMalloc(&m->w3p, sizeof(u64x2) * kW3Groups * 64) == hipSuccess &&
            hipMalloc(&m->bias, sizeof(float) * (2 * kHid + kOut)) == hipSuccess &&
            hipMalloc(&m->norm, sizeof(float) * kNorm) == hipSuccess &&
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
  // no standardisation until pairs or the caller set one
  std::vector<float> ident(kNorm, 0.0f);
  for (int k = 0; k < kIn; ++k) ident[kNormIsig + k] = 1.0f;
  for (int c = 0; c < kOut; ++c) ident[kNormSigT + c] = 1.0f;
  hipMemcpy(m->norm, ident.data(), sizeof(float) * kNorm, hipMemcpyHostToDevice);
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
  void* bufs[] = {m->w, m->m, m->v, m->g, m->w1p, m->w2p, m->w3p, m->bias, m->norm, m->feat,
                  m->targ, m->count, m->part, m->lpart, m->loss};
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

int bl_mlp_normalize(bl_mlp_t* m) {
  if (!m) return BL_FAILURE;
  return mlp_normalize(m);
}

int bl_mlp_norm(bl_mlp_t* m, float* get, const float* set) {
  if (!m) return BL_FAILURE;
  BL_HIP(hipStreamSynchronize(m->stream));
  if (get) BL_HIP(hipMemcpy(get, m->norm, sizeof(float) * kNorm, hipMemcpyDeviceToHost));
  if (set) {
    for (int c = 0; c < kOut; ++c)
      if (!(set[kNormSigT + c] > 0.0f)) {
        set_error("bl_mlp_norm: output standard deviations must be > 0");
        return BL_FAILURE;
      }
    BL_HIP(hipMemcpy(m->norm, set, sizeof(float) * kNorm, hipMemcpyHostToDevice));
    m->norm_set = true;
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
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, m->stream, packed(m),
                     (const float*)m->norm, (const float*)dx, n, dy);
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
  if (!m->norm_set && mlp_normalize(m)) return BL_FAILURE;
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
  bl_touch(h);
  if (!h) return BL_FAILURE;
  if (m && (!h->d_ring || h->relax_type == BL_RELAX_OFF)) {
    set_error("bl_relax_collect: set a plain relaxation first (bl_set_relaxation)");
    return BL_FAILURE;
  }
  if (m && stride < 1) {
    set_error("bl_relax_collect: stride >= 1");
    return BL_FAILURE;
  }
  BL_HIP(hipStreamSynchronize(h->stream));
  for (int k = 0; k < 2 && m; ++k)
    if (!h->d_cw[k]) BL_HIP(hipMalloc(&h->d_cw[k], sizeof(double) * std::max(h->sz.nnd, 1)));
  invalidate_graph(h);
  h->mlp_collect = m;
  h->collect_stride = m ? stride : 1;
  return BL_SUCCESS;
}

int bl_relax_learned(bl_handle_t* h, bl_mlp_t* m) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  if (m && (!h->d_ring || h->relax_type == BL_RELAX_OFF)) {
    set_error("bl_relax_learned: set the plain relaxation (the fallback) first");
    return BL_FAILURE;
  }
  BL_HIP(hipStreamSynchronize(h->stream));
  invalidate_graph(h);
  h->mlp_infer = m;
  if (m && !h->d_lflag) BL_HIP(hipMalloc(&h->d_lflag, sizeof(int32_t) * kLFlags));
  if (m && !h->d_lmark)
    BL_HIP(hipMalloc(&h->d_lmark, sizeof(int32_t) * std::max(h->sz.nnd, 1)));
  if (m && !h->d_ld1) BL_HIP(hipMalloc(&h->d_ld1, sizeof(float) * 2 * std::max(h->sz.nnd, 1)));
  // the counters start again with every network attached
  if (m) BL_HIP(hipMemset(h->d_lflag, 0, sizeof(int32_t) * kLFlags));
  return BL_SUCCESS;
}

int bl_set_learned_check(bl_handle_t* h, double max_aspect, int32_t relative) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  if (!(max_aspect >= 1.0) || (relative != 0 && relative != 1)) {
    set_error("bl_set_learned_check: max_aspect >= 1 and relative 0 or 1");
    return BL_FAILURE;
  }
  BL_HIP(hipStreamSynchronize(h->stream));
  invalidate_graph(h);
  h->learned_max_aspect = max_aspect;
  h->learned_relative = relative;
  return BL_SUCCESS;
}

int bl_relax_learned_stats(bl_handle_t* h, int32_t* steps, int32_t* guarded, int32_t* fallbacks,
                           int32_t* reasons) {
  if (!h) return BL_FAILURE;
  int32_t f[kLFlags] = {};
  if (h->d_lflag) {
    BL_HIP(hipStreamSynchronize(h->stream));
    BL_HIP(hipMemcpy(f, h->d_lflag, sizeof(int32_t) * kLFlags, hipMemcpyDeviceToHost));
  }
  // f[0], f[4] are the last step's, not yet added to f[2], f[5]
  if (steps) *steps = f[1];
  if (guarded) *guarded = f[5] + (f[1] > 0 && f[4] ? 1 : 0);
  if (fallbacks) *fallbacks = f[2] + (f[1] > 0 && f[0] ? 1 : 0);
  if (reasons) *reasons = f[3];
  return BL_SUCCESS;
}

// the features of every relaxed node at the current positions (nnd x 26,
// zero for nodes that are not relaxed)
int bl_relax_features(bl_handle_t* h, float* feat) {
  bl_touch(h);
  if (!h || !feat || !h->d_ring || !h->d_rlist || h->relax_type == BL_RELAX_OFF) {
    set_error("bl_relax_features: set a relaxation first (bl_set_relaxation)");
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
                     relax_args(h), (const int32_t*)h->d_rlist, nr, 1,
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
