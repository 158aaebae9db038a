// bl_diag.hip -- conservation tables (hydro_dr_shortprint) and the Sod L1
// check (check_dr_tests / check_kn_sod) on the device.
//
// Reductions have a fixed shape: each workgroup reduces its elements per
// flag with a wave64 butterfly and an LDS pass over its waves, then one
// workgroup adds the per-workgroup partials in workgroup order, one thread
// per (flag, quantity).  Results are reproducible run to run and differ from
// the reference's serial sums only in rounding.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstring>
#include <vector>

#include "bl_internal.h"
#include "bl_math.h"
#include "bl_mix.h"

namespace bl {

// quantities of one table row, in bl_flag_totals_t order; 0-4 are sums,
// 5/7 minima, 6/8 maxima
constexpr int kNq = 9;

__device__ __forceinline__ double q_identity(int q) {
  return (q == 5 || q == 7) ? DBL_MAX : ((q == 6 || q == 8) ? -DBL_MAX : 0.0);
}
__device__ __forceinline__ double q_combine(int q, double a, double b) {
  if (q == 5 || q == 7) return dmin(a, b);
  if (q == 6 || q == 8) return dmax(a, b);
  return a + b;
}

// one element's contribution to its flag's row (hydro_kn_calc_shortprint,
// hydro_kn_print.F90:81-107); cu/cv the corner velocities
__device__ __forceinline__ void sp_row(double dencut, double den, double en, double pre,
                                       double vol, double mass, const double wt[4],
                                       const double cu[4], const double cv[4], double r[kNq]) {
  r[0] = vol;
  r[1] = 0.0; r[2] = 0.0; r[3] = 0.0; r[4] = 0.0;
  r[5] = DBL_MAX; r[6] = -DBL_MAX; r[7] = DBL_MAX; r[8] = -DBL_MAX;
  if (mass > dencut * vol) {
    r[1] = mass;
    r[2] = en * mass;
    r[4] = mass * pre;
    r[5] = pre; r[6] = pre; r[7] = den; r[8] = den;
    double ke = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) ke = 0.5 * wt[c] * den * (cu[c] * cu[c] + cv[c] * cv[c]) + ke;
    r[3] = ke;
  }
}

// reduce r (this thread's contribution to flag `mine`, none if mine < 0)
// over the workgroup for every flag; thread 0 writes part[blockIdx][f][q].
// Every thread of the workgroup must call it.
__device__ void block_flag_reduce(int nflag, int mine, const double r[kNq], double* part) {
  __shared__ double sh[kBlock / 64][kNq];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int f = 0; f < nflag; ++f) {
    double v[kNq];
#pragma unroll
    for (int q = 0; q < kNq; ++q) v[q] = (mine == f) ? r[q] : q_identity(q);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int q = 0; q < kNq; ++q) v[q] = q_combine(q, v[q], __shfl_xor(v[q], off, 64));
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < kNq; ++q) sh[wid][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
        for (int q = 0; q < kNq; ++q) sh[0][q] = q_combine(q, sh[0][q], sh[w][q]);
      for (int q = 0; q < kNq; ++q) part[((size_t)blockIdx.x * nflag + f) * kNq + q] = sh[0][q];
    }
    __syncthreads();
  }
}

// element rows by flag (ielreg or ielmat; flag <= 0 skipped).  FROM_NODES:
// volume and corner weights from the node positions (the fused step keeps
// the geometry fields for the end of a batch); else from the fields
template <bool FROM_NODES>
__global__ void k_sp_elements(int nel, int nflag, double dencut, const int32_t* __restrict__ flag,
                              const int32_t* __restrict__ ielnd, const double* __restrict__ den,
                              const double* __restrict__ en, const double* __restrict__ pre,
                              const double* __restrict__ vol, const double* __restrict__ mass,
                              const double* __restrict__ cnwt, const double* __restrict__ ndx,
                              const double* __restrict__ ndy, const double* __restrict__ ndu,
                              const double* __restrict__ ndv, double* __restrict__ part,
                              const DevTimestep* ts) {
  if (ts && ts->skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double r[kNq];
  int mine = -1;
  if (e < nel && flag[e] > 0) {
    mine = flag[e] - 1;
    const int4 q = reinterpret_cast<const int4*>(ielnd)[e];
    const int nd[4] = {q.x - 1, q.y - 1, q.z - 1, q.w - 1};
    const double cu[4] = {ndu[nd[0]], ndu[nd[1]], ndu[nd[2]], ndu[nd[3]]};
    const double cv[4] = {ndv[nd[0]], ndv[nd[1]], ndv[nd[2]], ndv[nd[3]]};
    if (FROM_NODES) {
      const double x[4] = {ndx[nd[0]], ndx[nd[1]], ndx[nd[2]], ndx[nd[3]]};
      const double y[4] = {ndy[nd[0]], ndy[nd[1]], ndy[nd[2]], ndy[nd[3]]};
      Iso g;
      getiso(x, y, g);
      sp_row(dencut, den[e], en[e], pre[e], g.vol, mass[e], g.wt, cu, cv, r);
    } else {
      const double4 w = reinterpret_cast<const double4*>(cnwt)[e];
      const double wt[4] = {w.x, w.y, w.z, w.w};
      sp_row(dencut, den[e], en[e], pre[e], vol[e], mass[e], wt, cu, cv, r);
    }
  }
  block_flag_reduce(nflag, mine, r, part);
}

// component rows by icpmat; a component's corners take its cell's cnwt and
// corner velocities (utils_kn_mxgathercncp, hydro_dr_print.f90:96-118)
__global__ void k_sp_components(MixArgs M, int nflag, double dencut,
                                const int32_t* __restrict__ ielnd,
                                const double* __restrict__ cnwt, const double* __restrict__ ndu,
                                const double* __restrict__ ndv, const int32_t* __restrict__ cpel,
                                double* __restrict__ part) {
  const int icp = blockIdx.x * blockDim.x + threadIdx.x;
  double r[kNq];
  int mine = -1;
  const int e = icp < M.ncp ? cpel[icp] : -1;
  if (e >= 0 && M.icpmat[icp] > 0) {
    mine = M.icpmat[icp] - 1;
    const int4 q = reinterpret_cast<const int4*>(ielnd)[e];
    const int nd[4] = {q.x - 1, q.y - 1, q.z - 1, q.w - 1};
    const double4 w = reinterpret_cast<const double4*>(cnwt)[e];
    const double wt[4] = {w.x, w.y, w.z, w.w};
    const double cu[4] = {ndu[nd[0]], ndu[nd[1]], ndu[nd[2]], ndu[nd[3]]};
    const double cv[4] = {ndv[nd[0]], ndv[nd[1]], ndv[nd[2]], ndv[nd[3]]};
    sp_row(dencut, M.cpdensity[icp], M.cpenergy[icp], M.cppressure[icp], M.cpvolume[icp],
           M.cpmass[icp], wt, cu, cv, r);
  }
  block_flag_reduce(nflag, mine, r, part);
}

// element of each component, from the mixed-cell lists
__global__ void k_cp_element(MixArgs M, int32_t* __restrict__ cpel) {
  const int imx = blockIdx.x * blockDim.x + threadIdx.x;
  if (imx >= M.nmx) return;
  int icp = M.imxfcp[imx] - 1;
  for (int ii = 0; ii < M.imxncp[imx]; ++ii, ++icp) cpel[icp] = M.imxel[imx] - 1;
}

// sod() of check_kn_tests.f90:64-111
__device__ __forceinline__ void sod_exact(double xx, double time, double& r1, double& r2) {
  const double coeff = sqrt(1.4);
  r1 = -DBL_MAX; r2 = -DBL_MAX;
  if (xx <= -coeff * time) {
    r1 = 1.0; r2 = 2.5;
  } else if (xx <= -0.07027281256118326961 * time) {
    const double ss = xx / time;
    const double velocity = (5.0 / 6.0) * (ss + coeff);
    const double tt = ((velocity - ss) * (velocity - ss)) / 1.4;
    r1 = pow(tt, 2.5);
    r2 = r1 * tt / (0.4 * r1);
  } else if (xx <= 0.92745262004894994908 * time) {
    r1 = 0.42631942817849519385; r2 = 1.77760006942335264781;
  } else {
    if (time <= 28.53627624872492075028) {
      if (xx <= 1.75215573203017816370 * time) {
        r1 = 0.26557371170530706471; r2 = 2.85354088799095973738;
      } else {
        r1 = 0.125; r2 = 2.0;
      }
    } else if (time < 40.68191725689148697880) {
      if (xx <= (50.0 - (time - 28.53627624872492075028) * 1.01019363599101820444)) {
        r1 = 0.26557371170530706471; r2 = 2.85354088799095973738;
      } else {
        r1 = 0.50939531774381657731; r2 = 3.82996297061558546292;
      }
    }
  }
}

// check_kn_sod per element: (volume, |rho - rho_exact| V, |e - e_exact| V)
// at the centroid of geometry_kn_getcentroid (geometry_kn_centroid.f90:
// 29-55).  FROM_NODES: px/py are node positions gathered through ielnd and
// the volume is that of the node positions; else px/py are the corner
// arrays cnx/cny (4, nel) and the volume the elvolume field
template <bool FROM_NODES>
__global__ void k_l1(int nel, const int32_t* __restrict__ ielnd, const double* __restrict__ px,
                     const double* __restrict__ py, const double* __restrict__ vol,
                     const double* __restrict__ den, const double* __restrict__ en, double x0,
                     double scale, const double* __restrict__ ptime, double time,
                     double* __restrict__ part, const DevTimestep* ts) {
  if (ts && ts->skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double r[kNq];
  for (int q = 0; q < kNq; ++q) r[q] = q_identity(q);
  if (e < nel) {
    const double t = ptime ? *ptime : time;
    double x[4], y[4];
    if (FROM_NODES) {
      const int4 q4 = reinterpret_cast<const int4*>(ielnd)[e];
      const int nd[4] = {q4.x - 1, q4.y - 1, q4.z - 1, q4.w - 1};
#pragma unroll
      for (int c = 0; c < 4; ++c) { x[c] = px[nd[c]]; y[c] = py[nd[c]]; }
    } else {
      const double4 a = reinterpret_cast<const double4*>(px)[e];
      const double4 b = reinterpret_cast<const double4*>(py)[e];
      x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
      y[0] = b.x; y[1] = b.y; y[2] = b.z; y[3] = b.w;
    }
    double cx = 0.0, rv = 0.0;
#pragma unroll
    for (int i1 = 0; i1 < 4; ++i1) {
      const int i2 = (i1 + 1) & 3;
      const double rr = x[i1] * y[i2] - x[i2] * y[i1];
      rv = rv + rr;
      cx = cx + rr * (x[i1] + x[i2]);
    }
    rv = 0.5 * rv;
    cx = (rv > 0.0) ? cx / (6.0 * rv) : 0.0;
    double s1, s2;
    sod_exact(scale * (cx - x0), scale * t, s1, s2);
    double v;
    if (FROM_NODES) {
      Iso g;
      getiso(x, y, g);
      v = g.vol;
    } else {
      v = vol[e];
    }
    r[0] = v;
    r[1] = fabs(den[e] - s1) * v;
    r[2] = fabs(en[e] - s2) * v;
  }
  block_flag_reduce(1, 0, r, part);
}

// one column of the per-workgroup partials (part[b * stride + off], b <
// nb) reduced by one workgroup: each thread combines b = t, t + kBlock, ...
// in order, then a fixed pairwise tree over the threads in LDS.  The result
// depends only on nb and the data.  Every thread of the workgroup must call
// it; all get the result.
__device__ double column_reduce(int nb, size_t stride, size_t off, int q,
                                const double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int t = threadIdx.x;
  double a = q_identity(q);
  for (int b = t; b < nb; b += kBlock) a = q_combine(q, a, part[(size_t)b * stride + off]);
  sh[t] = a;
  __syncthreads();
#pragma unroll
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = q_combine(q, sh[t], sh[t + s]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// one workgroup per column i of an (nflag x kNq) table: out[i]
__global__ void __launch_bounds__(kBlock) k_flag_final(int nb, int nflag,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ out) {
  const int i = blockIdx.x;
  const double r = column_reduce(nb, (size_t)nflag * kNq, i, i % kNq, part);
  if (threadIdx.x == 0) out[i] = r;
}

// the columns of this step's record: nreg x kNq region columns from the
// partials of k_sp_elements (part), then the three L1 sums from k_l1's
// (part2): one workgroup per column, into col
__global__ void __launch_bounds__(kBlock) k_diag_columns(int nb, int nreg,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ part2,
                                                         double* __restrict__ col,
                                                         const DevTimestep* ts) {
  if (ts->skip) return;
  const int c = blockIdx.x, nt = nreg * kNq;
  const double r = c < nt ? column_reduce(nb, nt, c, c % kNq, part)
                          : column_reduce(nb, kNq, c - nt, 0, part2);
  if (threadIdx.x == 0) col[c] = r;
}

// the history record of this step, [nstep, time, region table (nreg x 9),
// L1 rho, L1 e], from the columns; then the record counter
__global__ void k_diag_record(int nreg, const double* __restrict__ col,
                              double* __restrict__ hist, int32_t* __restrict__ counter,
                              int maxrec, const DevTimestep* ts) {
  if (ts->skip) return;
  const int k = *counter;
  if (k >= maxrec) return;
  const int nt = nreg * kNq, rec_len = 2 + nt + 2;
  double* dst = hist + (size_t)k * rec_len;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) dst[2 + i] = col[i];
  if (threadIdx.x == 0) {
    dst[0] = (double)ts->nstep;
    dst[1] = ts->time;
    dst[2 + nt] = col[nt + 1] / col[nt];
    dst[3 + nt] = col[nt + 2] / col[nt];
  }
  __syncthreads();
  if (threadIdx.x == 0) *counter = k + 1;
}

}  // namespace bl

using namespace bl;

namespace {

// the partials buffer, at least n doubles; a new buffer invalidates the
// captured step, whose diagnostics kernels hold the old one
double* diag_part(bl_handle_t* h, size_t n) {
  if (h->diag_part_n < n) {
    hipStreamSynchronize(h->stream);
    invalidate_graph(h);
    if (h->d_diag_part) hipFree(h->d_diag_part);
    h->d_diag_part = nullptr;
    h->diag_part_n = 0;
    if (hipMalloc(&h->d_diag_part, sizeof(double) * n) != hipSuccess) {
      set_error("out of device memory for diagnostics");
      return nullptr;
    }
    h->diag_part_n = n;
  }
  return h->d_diag_part;
}

// the final pass over nb workgroups' partials into dout (nflag rows on the
// device), then to out on the host
int final_to_host(bl_handle_t* h, int nflag, unsigned nb, const double* part, double* dout,
                  double* out) {
  hipLaunchKernelGGL(k_flag_final, dim3(nflag * kNq), dim3(kBlock), 0, h->stream, (int)nb,
                     nflag, part, dout);
  hipError_t e0 = hipGetLastError();
  hipError_t e1 = hipMemcpyAsync(out, dout, sizeof(double) * kNq * nflag,
                                 hipMemcpyDeviceToHost, h->stream);
  hipError_t e2 = hipStreamSynchronize(h->stream);
  if (e0 != hipSuccess || e1 != hipSuccess || e2 != hipSuccess) {
    set_error("diagnostics reduction failed");
    return BL_FAILURE;
  }
  return BL_SUCCESS;
}

}  // namespace

namespace bl {

// the per-step record of the fused step from the node set the step wrote
int launch_diag(Handle* h, const double* ndx, const double* ndy, const double* ndu,
                const double* ndv, const double* rho, const double* en, const double* pre) {
  if (!h->diag_on) return BL_SUCCESS;
  if (!rho) rho = h->R(BL_ELDENSITY);
  if (!en) en = h->R(BL_ELENERGY);
  if (!pre) pre = h->R(BL_ELPRESSURE);
  const int nel = h->sz.nel, nreg = h->nreg;
  const unsigned nb = nblocks(nel);
  double* part = h->d_diag_part;  // sized by bl_set_diagnostics
  double* part2 = part + (size_t)nb * nreg * kNq;
  BL_LAUNCH(h, "diag_shortprint", k_sp_elements<true>, nb, kBlock, nel, nreg, h->cfg.dencut,
            (const int32_t*)h->I(BL_IELREG), (const int32_t*)h->I(BL_IELND), rho, en, pre,
            (const double*)nullptr, (const double*)h->R(BL_ELMASS), (const double*)nullptr, ndx,
            ndy, ndu, ndv, part, (const DevTimestep*)h->d_ts);
  BL_LAUNCH(h, "diag_l1", k_l1<true>, nb, kBlock, nel, (const int32_t*)h->I(BL_IELND), ndx, ndy,
            (const double*)nullptr, rho, en, h->diag_x0, h->diag_scale,
            (const double*)&h->d_ts->time, 0.0, part2, (const DevTimestep*)h->d_ts);
  double* col = part2 + (size_t)nb * kNq;
  BL_LAUNCH(h, "diag_columns", k_diag_columns, nreg * kNq + 3, kBlock, (int)nb, nreg,
            (const double*)part, (const double*)part2, col, (const DevTimestep*)h->d_ts);
  BL_LAUNCH(h, "diag_record", k_diag_record, 1, 64, nreg, (const double*)col,
            h->d_diag_hist, h->d_diag_count, h->diag_max, (const DevTimestep*)h->d_ts);
  return BL_SUCCESS;
}

}  // namespace bl

extern "C" {

int bl_shortprint(bl_handle_t* h, bl_flag_totals_t* region, bl_flag_totals_t* material) {
  bl_touch(h);
  if (!h || !region || !material) return BL_FAILURE;
  const int nel = h->sz.nel, nreg = h->sz.nreg, nmat = h->sz.nmat, ncp = h->sz.ncp;
  const unsigned nb = nblocks(nel), nbc = ncp > 0 ? nblocks(ncp) : 0;
  const int nf = std::max(std::max(nreg, nmat), 1);
  const size_t used = (size_t)(nb + nbc) * nf * kNq;
  double* part = diag_part(h, used + (size_t)nf * kNq);
  if (!part) return BL_FAILURE;
  double* dout = part + used;
  BL_LAUNCH(h, "shortprint", k_sp_elements<false>, nb, kBlock, nel, nreg, h->cfg.dencut,
            (const int32_t*)h->I(BL_IELREG), (const int32_t*)h->I(BL_IELND),
            (const double*)h->R(BL_ELDENSITY), (const double*)h->R(BL_ELENERGY),
            (const double*)h->R(BL_ELPRESSURE), (const double*)h->R(BL_ELVOLUME),
            (const double*)h->R(BL_ELMASS), (const double*)h->R(BL_CNWT),
            (const double*)nullptr, (const double*)nullptr, (const double*)h->R(BL_NDU),
            (const double*)h->R(BL_NDV), part, (const DevTimestep*)nullptr);
  if (final_to_host(h, nreg, nb, part, dout, reinterpret_cast<double*>(region)))
    return BL_FAILURE;
  BL_LAUNCH(h, "shortprint", k_sp_elements<false>, nb, kBlock, nel, nmat, h->cfg.dencut,
            (const int32_t*)h->I(BL_IELMAT), (const int32_t*)h->I(BL_IELND),
            (const double*)h->R(BL_ELDENSITY), (const double*)h->R(BL_ELENERGY),
            (const double*)h->R(BL_ELPRESSURE), (const double*)h->R(BL_ELVOLUME),
            (const double*)h->R(BL_ELMASS), (const double*)h->R(BL_CNWT),
            (const double*)nullptr, (const double*)nullptr, (const double*)h->R(BL_NDU),
            (const double*)h->R(BL_NDV), part, (const DevTimestep*)nullptr);
  if (ncp > 0) {
    int32_t* cpel = nullptr;
    BL_HIP(hipMalloc(&cpel, sizeof(int32_t) * ncp));
    // -1: a component in no mixed cell's range is left out
    BL_HIP(hipMemsetAsync(cpel, 0xff, sizeof(int32_t) * ncp, h->stream));
    MixArgs M = make_mix(h);
    hipLaunchKernelGGL(k_cp_element, dim3(nblocks(h->sz.nmx)), dim3(kBlock), 0, h->stream, M,
                       cpel);
    // the component partials follow the element partials: the final pass
    // adds them after the elements, as the reference's second call does
    hipLaunchKernelGGL(k_sp_components, dim3(nbc), dim3(kBlock), 0, h->stream, M, nmat,
                       h->cfg.dencut, (const int32_t*)h->I(BL_IELND),
                       (const double*)h->R(BL_CNWT), (const double*)h->R(BL_NDU),
                       (const double*)h->R(BL_NDV), (const int32_t*)cpel,
                       part + (size_t)nb * nmat * kNq);
    hipError_t e = hipGetLastError();
    hipStreamSynchronize(h->stream);
    hipFree(cpel);
    if (e != hipSuccess) { set_error("shortprint components failed"); return BL_FAILURE; }
  }
  return final_to_host(h, nmat, nb + nbc, part, dout, reinterpret_cast<double*>(material));
}

int bl_sod_l1(bl_handle_t* h, double time, double x0, double scale, double* l1) {
  bl_touch(h);
  if (!h || !l1) return BL_FAILURE;
  const int nel = h->sz.nel;
  const unsigned nb = nblocks(nel);
  double* part = diag_part(h, (size_t)(nb + 1) * kNq);
  if (!part) return BL_FAILURE;
  BL_LAUNCH(h, "sod_l1", k_l1<false>, nb, kBlock, nel, (const int32_t*)h->I(BL_IELND),
            (const double*)h->R(BL_CNX), (const double*)h->R(BL_CNY),
            (const double*)h->R(BL_ELVOLUME), (const double*)h->R(BL_ELDENSITY),
            (const double*)h->R(BL_ELENERGY), x0, scale, (const double*)nullptr, time, part,
            (const DevTimestep*)nullptr);
  double s[kNq];
  if (final_to_host(h, 1, nb, part, part + (size_t)nb * kNq, s)) return BL_FAILURE;
  l1[0] = s[1] / s[0];
  l1[1] = s[2] / s[0];
  return BL_SUCCESS;
}

int bl_set_diagnostics(bl_handle_t* h, int32_t on, double x0, double scale, int32_t maxrec) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  BL_HIP(hipStreamSynchronize(h->stream));
  invalidate_graph(h);
  h->diag_on = false;
  if (!on) return BL_SUCCESS;
  if (maxrec <= 0 || h->nreg <= 0) {
    set_error("bl_set_diagnostics: maxrec and the region count must be positive");
    return BL_FAILURE;
  }
  const int nreg = h->nreg;
  const unsigned nb = nblocks(h->sz.nel);
  if (!diag_part(h, (size_t)nb * (nreg + 1) * kNq + (size_t)nreg * kNq + 3)) return BL_FAILURE;
  if (h->d_diag_hist) hipFree(h->d_diag_hist);
  h->d_diag_hist = nullptr;
  const size_t rec = 2 + kNq * (size_t)nreg + 2;
  BL_HIP(hipMalloc(&h->d_diag_hist, sizeof(double) * rec * maxrec));
  if (!h->d_diag_count) BL_HIP(hipMalloc(&h->d_diag_count, sizeof(int32_t)));
  BL_HIP(hipMemsetAsync(h->d_diag_count, 0, sizeof(int32_t), h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  h->diag_on = true;
  h->diag_x0 = x0;
  h->diag_scale = scale;
  h->diag_max = maxrec;
  return BL_SUCCESS;
}

int bl_diagnostics(bl_handle_t* h, int32_t maxrec, double* out, int32_t* nrec) {
  bl_touch(h);
  if (!h || !nrec) return BL_FAILURE;
  *nrec = 0;
  if (!h->d_diag_count || !h->d_diag_hist) return BL_SUCCESS;
  int32_t n = 0;
  BL_HIP(hipMemcpyAsync(&n, h->d_diag_count, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  n = std::min(n, std::min(maxrec, h->diag_max));
  const size_t rec = 2 + kNq * (size_t)h->nreg + 2;
  if (n > 0 && out)
    BL_HIP(hipMemcpy(out, h->d_diag_hist, sizeof(double) * rec * n, hipMemcpyDeviceToHost));
  *nrec = n;
  return BL_SUCCESS;
}

}  // extern "C"
