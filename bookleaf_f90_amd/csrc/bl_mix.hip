// bl_mix.hip -- component terms of the hydro drivers for mixed cells.
//
// One thread per mixed cell walks the cell's contiguous components (the
// element-to-component gathers and the component-to-element averages of
// utils_kn_gather.f90:53-154 and utils_kn_average.f90:29-107); purely
// per-component work uses one thread per component.  Launched only when
// ncp > 0.
#include <hip/hip_runtime.h>

#include "bl_internal.h"
#include "bl_math.h"
#include "bl_mix.h"

namespace bl {

#define MIX_GATE(ts) \
  if ((ts) && ((ts)->skip || !(ts)->ale_on)) return

// geometry_dr_volume.f90:66-79: element iso terms to the components
// (utils_kn_mxgather4), component volume 4(a1 b3 - a3 b1) times frvolume
__global__ void k_mx_geometry(MixArgs M, const double* __restrict__ a1,
                              const double* __restrict__ a3, const double* __restrict__ b1,
                              const double* __restrict__ b3, const DevTimestep* ts) {
  MIX_GATE(ts);
  const int imx = xcd_index();
  if (imx >= M.nmx) return;
  const int iel = M.imxel[imx] - 1;
  const double w1 = a1[iel], w2 = a3[iel], w3 = b1[iel], w4 = b3[iel];
  int icp = M.imxfcp[imx] - 1;
  for (int ii = 0; ii < M.imxncp[imx]; ++ii, ++icp) {
    M.cpa1[icp] = w1; M.cpa3[icp] = w2; M.cpb1[icp] = w3; M.cpb3[icp] = w4;
    M.cpvolume[icp] = (4.0 * (w1 * w4 - w2 * w3)) * M.frvolume[icp];
  }
}

// the component part of ale_dr_hydroupdate (ale_dr_update.f90:33-51):
// geometry_dr_volume.f90:66-79, density_dr_getdensity.f90:40-43 and
// eos_dr_geteos.f90:50-55 per mixed cell
__global__ void k_mx_hydroupdate(MixArgs M, const int32_t* __restrict__ ielnd,
                                 const double* __restrict__ ndx, const double* __restrict__ ndy,
                                 const int32_t* __restrict__ eos_itype,
                                 const double* __restrict__ eos_param, double pcut, double ccut,
                                 const DevTimestep* ts) {
  MIX_GATE(ts);
  const int imx = xcd_index();
  if (imx >= M.nmx) return;
  mix_hydroupdate_cell(M, imx, ielnd, ndx, ndy, eos_itype, eos_param, pcut, ccut);
}

// density_dr_getdensity.f90:40-43
__global__ void k_mx_density(MixArgs M, const DevTimestep* ts) {
  MIX_GATE(ts);
  const int icp = xcd_index();
  if (icp >= M.ncp) return;
  M.cpdensity[icp] = M.cpmass[icp] / M.cpvolume[icp];
}

// eos_dr_geteos.f90:50-55 with icpmat
__global__ void k_mx_eos(MixArgs M, const int32_t* __restrict__ eos_itype,
                         const double* __restrict__ eos_param, double pcut, double ccut,
                         const DevTimestep* ts) {
  MIX_GATE(ts);
  const int icp = xcd_index();
  if (icp >= M.ncp) return;
  int mat = M.icpmat[icp];
  mat = mat > 0 ? mat : 0;
  const int et = eos_itype[mat];
  const double* ep = eos_param + BL_NEOSPARAM * mat;
  const double r = M.cpdensity[icp], e = M.cpenergy[icp];
  M.cppressure[icp] = eos_pressure(et, ep, r, e, pcut);
  M.cpcs2[icp] = eos_cs2(et, ep, r, e, pcut, ccut);
}

// init_dr_mass, init_dr_utils.f90:121-130: cpmass = cpdensity*cpvolume
// (init_kn_mxmass), elmass gathered into rcpscratch11, frmass = cpmass/that
__global__ void k_mx_initmass(MixArgs M, const double* __restrict__ elmass,
                              double* __restrict__ rcps11) {
  const int imx = xcd_index();
  if (imx >= M.nmx) return;
  const double w = elmass[M.imxel[imx] - 1];
  int icp = M.imxfcp[imx] - 1;
  for (int ii = 0; ii < M.imxncp[imx]; ++ii, ++icp) {
    const double cm = M.cpdensity[icp] * M.cpvolume[icp];
    M.cpmass[icp] = cm;
    rcps11[icp] = w;
    M.frmass[icp] = cm / w;
  }
}

// hydro_kn_viscinit over the components (hydro_dr_init.f90:48-51)
__global__ void k_mx_viscinit(int ncp, double* __restrict__ cpvisc,
                              double* __restrict__ cpviscx, double* __restrict__ cpviscy) {
  const int icp = xcd_index();
  if (icp >= ncp) return;
  cpvisc[icp] = 0.0;
  const double z[4] = {0.0, 0.0, 0.0, 0.0};
  st4c(cpviscx, icp, z);
  st4c(cpviscy, icp, z);
}

// hydro_kn_getforcep + hydro_kn_getforceq over the components
__global__ void k_mx_force(MixArgs M) {
  const int icp = xcd_index();
  if (icp >= M.ncp) return;
  double fx[4], fy[4];
  mix_force(M, icp, M.cppressure[icp], M.cpa1[icp], M.cpa3[icp], M.cpb1[icp], M.cpb3[icp],
            fx, fy);
  st4c(M.cpfx, icp, fx);
  st4c(M.cpfy, icp, fy);
}

// hydro_dr_getenergy.f90:47-62: cpu/cpv = frvolume * element corner
// velocity (utils_kn_mxgathercnav) into rscratch23/24, then the component
// energy update (hydro_kn_getenergy)
__global__ void k_mx_energy(MixArgs M, const double* __restrict__ cnu,
                            const double* __restrict__ cnv, double* __restrict__ cpu,
                            double* __restrict__ cpv, double dts, double zerocut) {
  const int imx = xcd_index();
  if (imx >= M.nmx) return;
  const int iel = M.imxel[imx] - 1;
  double w1[4], w2[4];
  ld4c(cnu, iel, w1);
  ld4c(cnv, iel, w2);
  int icp = M.imxfcp[imx] - 1;
  for (int ii = 0; ii < M.imxncp[imx]; ++ii, ++icp) {
    const double f = M.frvolume[icp];
    double cu[4], cv[4], fx[4], fy[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { cu[j] = f * w1[j]; cv[j] = f * w2[j]; }
    st4c(cpu, icp, cu);
    st4c(cpv, icp, cv);
    ld4c(M.cpfx, icp, fx);
    ld4c(M.cpfy, icp, fy);
    M.cpenergy[icp] = energy_update(M.cpenergy[icp], M.cpmass[icp], fx, fy, cu, cv, dts, zerocut);
  }
}

// hydro_dr_initacceleration: element corner forces of a mixed cell become
// the frvolume average of the component forces (utils_kn_average22)
__global__ void k_mx_average_force(MixArgs M, double* __restrict__ cnfx,
                                   double* __restrict__ cnfy) {
  const int imx = xcd_index();
  if (imx >= M.nmx) return;
  double w1[4] = {0.0, 0.0, 0.0, 0.0}, w2[4] = {0.0, 0.0, 0.0, 0.0};
  int icp = M.imxfcp[imx] - 1;
  for (int ii = 0; ii < M.imxncp[imx]; ++ii, ++icp) {
    const double f = M.frvolume[icp];
    double fx[4], fy[4];
    ld4c(M.cpfx, icp, fx);
    ld4c(M.cpfy, icp, fy);
#pragma unroll
    for (int j = 0; j < 4; ++j) { w1[j] = w1[j] + f * fx[j]; w2[j] = w2[j] + f * fy[j]; }
  }
  const int iel = M.imxel[imx] - 1;
  st4c(cnfx, iel, w1);
  st4c(cnfy, iel, w2);
}

// cpu/cpv of the corrector's getenergy from the node ubar (the corrector's
// cnu/cnv are ubar gathered to the corners, hydro_dr_applyacceleration)
__global__ void k_mx_cpu_overlay(MixArgs M, const int32_t* __restrict__ ielnd,
                                 const double* __restrict__ ubar,
                                 const double* __restrict__ vbar, double* __restrict__ cpu,
                                 double* __restrict__ cpv, const DevTimestep* ts) {
  if (ts && ts->skip) return;
  const int imx = xcd_index();
  if (imx >= M.nmx) return;
  const int iel = M.imxel[imx] - 1;
  const int4 q = reinterpret_cast<const int4*>(ielnd)[iel];
  const double w1[4] = {ubar[q.x - 1], ubar[q.y - 1], ubar[q.z - 1], ubar[q.w - 1]};
  const double w2[4] = {vbar[q.x - 1], vbar[q.y - 1], vbar[q.z - 1], vbar[q.w - 1]};
  int icp = M.imxfcp[imx] - 1;
  for (int ii = 0; ii < M.imxncp[imx]; ++ii, ++icp) {
    const double f = M.frvolume[icp];
    double cu[4], cv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { cu[j] = f * w1[j]; cv[j] = f * w2[j]; }
    st4c(cpu, icp, cu);
    st4c(cpv, icp, cv);
  }
}

MixArgs make_mix(Handle* h) {
  MixArgs M{};
  M.ncp = h->sz.ncp;
  M.nmx = h->sz.nmx;
  M.imxel = h->I(BL_IMXEL); M.imxfcp = h->I(BL_IMXFCP); M.imxncp = h->I(BL_IMXNCP);
  M.icpmat = h->I(BL_ICPMAT);
  M.frvolume = h->R(BL_FRVOLUME); M.frmass = h->R(BL_FRMASS);
  M.cpdensity = h->R(BL_CPDENSITY); M.cpenergy = h->R(BL_CPENERGY);
  M.cppressure = h->R(BL_CPPRESSURE); M.cpcs2 = h->R(BL_CPCS2);
  M.cpvolume = h->R(BL_CPVOLUME); M.cpmass = h->R(BL_CPMASS);
  M.cpvisc = h->R(BL_CPVISC); M.cpviscx = h->R(BL_CPVISCX); M.cpviscy = h->R(BL_CPVISCY);
  M.cpa1 = h->R(BL_CPA1); M.cpa3 = h->R(BL_CPA3); M.cpb1 = h->R(BL_CPB1); M.cpb3 = h->R(BL_CPB3);
  M.cpfx = h->R(BL_RCPSCRATCH21); M.cpfy = h->R(BL_RCPSCRATCH22);
  return M;
}

int mix_geometry(Handle* h, const DevTimestep* ts) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_geometry", k_mx_geometry, nblocks(h->sz.nmx), kBlock, make_mix(h),
            (const double*)h->R(BL_A1), (const double*)h->R(BL_A3),
            (const double*)h->R(BL_B1), (const double*)h->R(BL_B3), ts);
  return BL_SUCCESS;
}

int mix_density(Handle* h, const DevTimestep* ts) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_density", k_mx_density, nblocks(h->sz.ncp), kBlock, make_mix(h), ts);
  return BL_SUCCESS;
}

int mix_eos(Handle* h, const DevTimestep* ts) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_eos", k_mx_eos, nblocks(h->sz.ncp), kBlock, make_mix(h),
            (const int32_t*)h->d_eos_itype, (const double*)h->d_eos_param, h->cfg.pcut,
            h->cfg.ccut, ts);
  return BL_SUCCESS;
}

int mix_hydroupdate(Handle* h, const double* x, const double* y, const DevTimestep* ts) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_hydroupdate", k_mx_hydroupdate, nblocks(h->sz.nmx), kBlock, make_mix(h),
            (const int32_t*)h->I(BL_IELND), x, y, (const int32_t*)h->d_eos_itype,
            (const double*)h->d_eos_param, h->cfg.pcut, h->cfg.ccut, ts);
  return BL_SUCCESS;
}

int mix_initmass(Handle* h) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_initmass", k_mx_initmass, nblocks(h->sz.nmx), kBlock, make_mix(h),
            (const double*)h->R(BL_ELMASS), h->R(BL_RCPSCRATCH11));
  return BL_SUCCESS;
}

int mix_viscinit(Handle* h) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_viscinit", k_mx_viscinit, nblocks(h->sz.ncp), kBlock, h->sz.ncp,
            h->R(BL_CPVISC), h->R(BL_CPVISCX), h->R(BL_CPVISCY));
  return BL_SUCCESS;
}

int mix_force(Handle* h) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_force", k_mx_force, nblocks(h->sz.ncp), kBlock, make_mix(h));
  return BL_SUCCESS;
}

int mix_energy(Handle* h, double dts) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_energy", k_mx_energy, nblocks(h->sz.nmx), kBlock, make_mix(h),
            (const double*)h->R(BL_RSCRATCH21), (const double*)h->R(BL_RSCRATCH22),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24), dts, h->cfg.zerocut);
  return BL_SUCCESS;
}

int mix_average_force(Handle* h) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_average_force", k_mx_average_force, nblocks(h->sz.nmx), kBlock, make_mix(h),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24));
  return BL_SUCCESS;
}

int mix_cpu_overlay(Handle* h, const DevTimestep* ts) {
  if (h->sz.ncp <= 0) return BL_SUCCESS;
  BL_LAUNCH(h, "mx_cpu_overlay", k_mx_cpu_overlay, nblocks(h->sz.nmx), kBlock, make_mix(h),
            (const int32_t*)h->I(BL_IELND), (const double*)h->R(BL_RSCRATCH14),
            (const double*)h->R(BL_RSCRATCH15), h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24), ts);
  return BL_SUCCESS;
}

}  // namespace bl
