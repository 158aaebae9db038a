// bl_step.hip -- whole-step runners and conservation totals.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstring>

#include "bl_internal.h"
#include "bl_math.h"
#include "bl_step.h"

namespace bl {

static void fill_err(bl_error_t* err, int ierr, int iout, const char* msg) {
  if (!err) return;
  err->ierr = ierr;
  err->iout = iout;
  std::memset(err->serr, ' ', BL_SLEN);
  size_t n = std::strlen(msg);
  std::memcpy(err->serr, msg, n < (size_t)BL_SLEN ? n : (size_t)BL_SLEN);
}

// per-workgroup partial sums of hydro_kn_calc_shortprint's totals
// (hydro_kn_print.F90:56-110) for the whole mesh
__global__ void k_totals(int nel, double dencut, const int32_t* __restrict__ ielnd,
                         const double* __restrict__ rho, const double* __restrict__ en,
                         const double* __restrict__ pre, const double* __restrict__ vol,
                         const double* __restrict__ mass, const double* __restrict__ cnwt,
                         const double* __restrict__ ndu, const double* __restrict__ ndv,
                         double* __restrict__ part) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0, m = 0.0, ie = 0.0, ke = 0.0;
  double pmn = DBL_MAX, pmx = -DBL_MAX, dmn = DBL_MAX, dmx = -DBL_MAX;
  if (e < nel) {
    double V = vol[e];
    v = V;
    double M = mass[e];
    if (M > dencut * V) {
      m = M;
      ie = en[e] * M;
      double p = pre[e], d = rho[e];
      pmn = p; pmx = p; dmn = d; dmx = d;
      int4 nd = reinterpret_cast<const int4*>(ielnd)[e];
      double4 w = reinterpret_cast<const double4*>(cnwt)[e];
      int n4[4] = {nd.x - 1, nd.y - 1, nd.z - 1, nd.w - 1};
      double w4[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double u = ndu[n4[c]], vv = ndv[n4[c]];
        ke = 0.5 * w4[c] * d * (u * u + vv * vv) + ke;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off, 64);
    m += __shfl_xor(m, off, 64);
    ie += __shfl_xor(ie, off, 64);
    ke += __shfl_xor(ke, off, 64);
    pmn = dmin(pmn, __shfl_xor(pmn, off, 64));
    pmx = dmax(pmx, __shfl_xor(pmx, off, 64));
    dmn = dmin(dmn, __shfl_xor(dmn, off, 64));
    dmx = dmax(dmx, __shfl_xor(dmx, off, 64));
  }
  __shared__ double s[kBlock / 64][8];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s[wid][0] = v; s[wid][1] = m; s[wid][2] = ie; s[wid][3] = ke;
    s[wid][4] = pmn; s[wid][5] = pmx; s[wid][6] = dmn; s[wid][7] = dmx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      for (int k = 0; k < 4; ++k) s[0][k] += s[w][k];
      s[0][4] = dmin(s[0][4], s[w][4]); s[0][5] = dmax(s[0][5], s[w][5]);
      s[0][6] = dmin(s[0][6], s[w][6]); s[0][7] = dmax(s[0][7], s[w][7]);
    }
    for (int k = 0; k < 8; ++k) part[8 * blockIdx.x + k] = s[0][k];
  }
}

}  // namespace bl

using namespace bl;

extern "C" {

int bl_totals(bl_handle_t* h, bl_totals_t* out) {
  bl_touch(h);
  if (!h || !out) return BL_FAILURE;
  int nel = h->sz.nel;
  unsigned nb = nblocks(nel);
  double* part = nullptr;
  BL_HIP(hipMalloc(&part, sizeof(double) * 8 * nb));
  BL_LAUNCH(h, "totals", k_totals, nb, kBlock, nel, h->cfg.dencut, h->I(BL_IELND),
            h->R(BL_ELDENSITY), h->R(BL_ELENERGY), h->R(BL_ELPRESSURE), h->R(BL_ELVOLUME),
            h->R(BL_ELMASS), h->R(BL_CNWT), h->R(BL_NDU), h->R(BL_NDV), part);
  std::vector<double> hp(8 * (size_t)nb);
  hipError_t e1 = hipMemcpyAsync(hp.data(), part, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, h->stream);
  hipError_t e2 = hipStreamSynchronize(h->stream);
  hipFree(part);
  if (e1 != hipSuccess || e2 != hipSuccess) { set_error("bl_totals: copy failed"); return BL_FAILURE; }
  double t[8] = {0, 0, 0, 0, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
  for (unsigned b = 0; b < nb; ++b) {
    for (int k = 0; k < 4; ++k) t[k] += hp[8 * b + k];
    t[4] = std::min(t[4], hp[8 * b + 4]); t[5] = std::max(t[5], hp[8 * b + 5]);
    t[6] = std::min(t[6], hp[8 * b + 6]); t[7] = std::max(t[7], hp[8 * b + 7]);
  }
  out->volume = t[0]; out->mass = t[1]; out->ie = t[2]; out->ke = t[3];
  out->pressure_min = t[4]; out->pressure_max = t[5];
  out->density_min = t[6]; out->density_max = t[7];
  return BL_SUCCESS;
}

// one step through the driver entry points, host dt selection
static int host_step(bl_handle_t* h, bl_timestep_t* ts, bl_error_t* err) {
  bl_dt_t cfl, div;
  if (bl_hydro_getdt(h, &cfl, &div, err) != BL_SUCCESS) return BL_FAILURE;
  if (select_dt(h->cfg, ts, cfl, div) != BL_SUCCESS) {
    fill_err(err, BL_FAILURE, BL_HALT_ALL, "ERROR: dt < dt_min");
    set_error("dt < dt_min");
    return BL_FAILURE;
  }
  ts->zcorrector = 0;
  ts->dts = 0.5 * ts->dt;
  if (bl_hydro_setpredictor(h) || bl_hydro_getforce(h, ts)) return BL_FAILURE;
  if (bl_hydro_getstate(h, ts, err)) return BL_FAILURE;
  if (bl_hydro_getenergy(h, ts) || bl_hydro_geteos(h)) return BL_FAILURE;
  ts->zcorrector = 1;
  ts->dts = ts->dt;
  if (bl_hydro_setcorrector(h) || bl_hydro_getforce(h, ts) || bl_hydro_initacceleration(h) ||
      bl_hydro_scatteracceleration(h) || bl_hydro_getacceleration(h) ||
      bl_hydro_applyacceleration(h, ts))
    return BL_FAILURE;
  if (bl_hydro_getstate(h, ts, err)) return BL_FAILURE;
  if (bl_hydro_getenergy(h, ts) || bl_hydro_geteos(h)) return BL_FAILURE;
  if (bl_ale_active(h, ts)) {
    if (bl_ale_getmeshstatus(h) || bl_ale_getfluxvolume(h, ts)) return BL_FAILURE;
    if (bl_ale_advect(h, ts, err) || bl_ale_hydroupdate(h, err)) return BL_FAILURE;
  }
  fill_err(err, BL_SUCCESS, 0, "");
  return BL_SUCCESS;
}

int bl_run(bl_handle_t* h, int32_t nsteps, int32_t stop_at_time_end, bl_timestep_t* ts,
           double* dt_hist, int32_t* idt_hist, int32_t* ctl_hist, int32_t* steps_taken,
           bl_error_t* err) {
  bl_touch(h);
  if (!h || !ts) return BL_FAILURE;
  int taken = 0, rc = BL_SUCCESS;
  fill_err(err, BL_SUCCESS, 0, "");
  for (int k = 0; k < nsteps; ++k) {
    if (stop_at_time_end && ts->time >= h->cfg.time_end) break;
    rc = host_step(h, ts, err);
    if (rc == BL_SUCCESS && h->diag_on) {
      // the record of this step, as the graph runner writes it
      DevTimestep* d = h->h_ts;
      std::memset(d, 0, sizeof(DevTimestep));
      d->time = ts->time; d->dt = ts->dt; d->nstep = ts->nstep;
      BL_HIP(hipMemcpyAsync(h->d_ts, d, sizeof(DevTimestep), hipMemcpyHostToDevice, h->stream));
      BL_HIP(hipStreamSynchronize(h->stream));
      rc = launch_diag(h, h->R(BL_NDX), h->R(BL_NDY), h->R(BL_NDU), h->R(BL_NDV));
    }
    if (dt_hist) dt_hist[k] = ts->dt;
    if (idt_hist) idt_hist[k] = ts->idtel;
    if (ctl_hist) ctl_hist[k] = ts->sdt;
    ++taken;
    if (rc != BL_SUCCESS) break;
  }
  if (steps_taken) *steps_taken = taken;
  if (rc == BL_SUCCESS && hipStreamSynchronize(h->stream) != hipSuccess) return BL_FAILURE;
  return rc;
}

}  // extern "C"
