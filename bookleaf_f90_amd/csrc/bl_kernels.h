// bl_kernels.h -- deterministic workgroup reductions used by the dt controls.
//
// hydro_kn_getdtcfl picks the FIRST element attaining the minimum of
// rscratch11 (strict '<' scan, hydro_kn_getdt.f90:68-72); hydro_kn_getdtdiv
// picks the FIRST element attaining the maximum, and only if it exceeds TINY
// (strict '>' scan from TINY, hydro_kn_getdt.f90:104-117).  Both are total
// orders on (value, index), so a tree reduction in any order returns the
// same element as the serial scan: wave64 shuffles, then LDS across the
// waves of the workgroup, then the last workgroup over the partials.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>

#include "bl_internal.h"

namespace bl {

__device__ __forceinline__ void better_min(double& v, int& i, double v2, int i2) {
  if (v2 < v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}
__device__ __forceinline__ void better_max(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

template <bool IS_MIN>
__device__ __forceinline__ void wave_reduce(double& v, int& i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double v2 = __shfl_xor(v, off, 64);
    int i2 = __shfl_xor(i, off, 64);
    if (IS_MIN) better_min(v, i, v2, i2);
    else better_max(v, i, v2, i2);
  }
}

// reduce over the workgroup; thread 0 writes the result to (*outv, *outi)
template <bool IS_MIN>
__device__ __forceinline__ void block_reduce_at(double v, int i, double* outv, int32_t* outi) {
  __shared__ double sv[kBlock / 64];
  __shared__ int si[kBlock / 64];
  wave_reduce<IS_MIN>(v, i);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sv[wid] = v; si[wid] = i; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bv = sv[0];
    int bi = si[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      if (IS_MIN) better_min(bv, bi, sv[w], si[w]);
      else better_max(bv, bi, sv[w], si[w]);
    }
    *outv = bv;
    *outi = bi;
  }
}

// the DIV maximum (dv, di) and the CFL minimum (cv, ci) over the workgroup
// together: one pass of shuffles for both, one barrier; thread 0 writes them
__device__ __forceinline__ void block_reduce_div_cfl(double dv, int di, double cv, int ci,
                                                     double* out_dv, int32_t* out_di,
                                                     double* out_cv, int32_t* out_ci) {
  __shared__ double sv[2][kBlock / 64];
  __shared__ int si[2][kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double dv2 = __shfl_xor(dv, off, 64), cv2 = __shfl_xor(cv, off, 64);
    const int di2 = __shfl_xor(di, off, 64), ci2 = __shfl_xor(ci, off, 64);
    better_max(dv, di, dv2, di2);
    better_min(cv, ci, cv2, ci2);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    sv[0][wid] = dv; si[0][wid] = di;
    sv[1][wid] = cv; si[1][wid] = ci;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      better_max(dv, di, sv[0][w], si[0][w]);
      better_min(cv, ci, sv[1][w], si[1][w]);
    }
    *out_dv = dv; *out_di = di;
    *out_cv = cv; *out_ci = ci;
  }
}

// the partial of this workgroup at part[blockIdx.x]
template <bool IS_MIN>
__device__ __forceinline__ void block_reduce_store(double v, int i, double* part, int32_t* parti) {
  block_reduce_at<IS_MIN>(v, i, part + blockIdx.x, parti + blockIdx.x);
}

__device__ __forceinline__ void block_argmin_first(double v, int i, double* part, int32_t* parti) {
  if (!(v == v)) v = DBL_MAX;  // a NaN never wins the strict '<' scan
  block_reduce_store<true>(v, i, part, parti);
}
__device__ __forceinline__ void block_argmax_first(double v, int i, double* part, int32_t* parti) {
  if (!(v == v)) v = -1.0;  // a NaN never wins the strict '>' scan
  block_reduce_store<false>(v, i, part, parti);
}

// run by one workgroup of up to kFinalBlock threads: DIV partials in
// part[0,na), CFL partials in part[2na, 2na+nc); both in one pass so the
// loads of a thread are independent
constexpr int kFinalBlock = 1024;
// this thread's share of the DIV partials (part, parti)[0, na) and the CFL
// partials (pc, qc)[0, nc): DIV maximum (va, xa), CFL minimum (vc, xc)
__device__ __forceinline__ void final_dt_gather(int na, const double* part, const int32_t* parti,
                                                int nc, const double* pc, const int32_t* qc,
                                                double& va, int& xa, double& vc, int& xc) {
  va = -1.0; vc = DBL_MAX;
  xa = 0x7fffffff; xc = 0x7fffffff;
  const int nmax = na > nc ? na : nc;
  if (na < 1 || nc < 1) {  // (not a case the step launches)
    for (int k = threadIdx.x; k < nmax; k += blockDim.x) {
      if (k < na) better_max(va, xa, part[k], parti[k]);
      if (k < nc) better_min(vc, xc, pc[k], qc[k]);
    }
    return;
  }
  // kPer entries of each control per thread and pass, all loaded (from
  // clamped indices, so with no branch) before any is compared
  constexpr int kPer = 4;
  const int stride = (int)blockDim.x;
  for (int b = threadIdx.x; b < nmax; b += kPer * stride) {
    double a[kPer], c[kPer];
    int ai[kPer], ci[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int k = b + r * stride;
      const int ka = k < na ? k : na - 1, kc = k < nc ? k : nc - 1;
      a[r] = part[ka]; ai[r] = parti[ka];
      c[r] = pc[kc]; ci[r] = qc[kc];
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int k = b + r * stride;
      if (k < na) better_max(va, xa, a[r], ai[r]);
      if (k < nc) better_min(vc, xc, c[r], ci[r]);
    }
  }
}
// the workgroup's combination of the threads' shares into res
__device__ __forceinline__ void final_dt_combine(double va, int xa, double vc, int xc,
                                                 DtResult* res) {
  __shared__ double sa[kFinalBlock / 64], sc[kFinalBlock / 64];
  __shared__ int ia[kFinalBlock / 64], ic[kFinalBlock / 64];
  wave_reduce<false>(va, xa);
  wave_reduce<true>(vc, xc);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sa[wid] = va; ia[wid] = xa; sc[wid] = vc; ic[wid] = xc; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      better_max(sa[0], ia[0], sa[w], ia[w]);
      better_min(sc[0], ic[0], sc[w], ic[w]);
    }
    if (sa[0] > DBL_MIN) { res->div_max = sa[0]; res->div_idx = ia[0]; }
    else { res->div_max = DBL_MIN; res->div_idx = 0; }
    res->cfl_min = sc[0];
    res->cfl_idx = ic[0];
  }
  __syncthreads();
}
// DIV partials (pa, qa)[0, na), CFL partials (pc, qc)[0, nc)
__device__ __forceinline__ void final_dt_reduce(int na, const double* part, const int32_t* parti,
                                                int nc, const double* pc, const int32_t* qc,
                                                DtResult* res) {
  double va, vc;
  int xa, xc;
  final_dt_gather(na, part, parti, nc, pc, qc, va, xa, vc, xc);
  final_dt_combine(va, xa, vc, xc, res);
}
__device__ __forceinline__ void final_dt_reduce(int na, int nc, const double* part,
                                                const int32_t* parti, DtResult* res) {
  final_dt_reduce(na, part, parti, nc, part + 2 * na, parti + 2 * na, res);
}

}  // namespace bl
