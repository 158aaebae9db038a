// bl_mix.h -- mixed cells (ncp > 0): component fields and the component
// terms of the hydro drivers.
//
// A mixed element iel has ielmat(iel) = -imx; its components are the
// contiguous range imxfcp(imx) .. imxfcp(imx)+imxncp(imx)-1 of the flattened
// lists (mix_dr_flatten, mix_dr_list.f90).  Every sum over a cell's
// components runs in that order, as utils_kn_average* do
// (utils_kn_average.f90:29-107), so the element averages are bit-identical.
// Component corner arrays are (4, mcp), component j of component icp at
// 4*icp + j.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bl_internal.h"
#include "bl_math.h"

namespace bl {

struct MixArgs {
  int ncp, nmx;
  const int32_t* imxel;   // 1-based element of each mixed cell
  const int32_t* imxfcp;  // 1-based first component
  const int32_t* imxncp;
  const int32_t* icpmat;
  const double* frvolume;
  double* frmass;
  double* cpdensity;
  double* cpenergy;
  double* cppressure;
  double* cpcs2;
  double* cpvolume;
  double* cpmass;
  const double* cpvisc;
  const double* cpviscx;  // (4, mcp)
  const double* cpviscy;
  double* cpa1;
  double* cpa3;
  double* cpb1;
  double* cpb3;
  double* cpfx;  // rcpscratch21 (lagstep cpfx), (4, mcp)
  double* cpfy;  // rcpscratch22
};

MixArgs make_mix(Handle* h);

// component forces of hydro_dr_getforce (hydro_dr_getforce.f90:52-66):
// pressure force with the component pressure and iso terms, plus cpvisc
__device__ __forceinline__ void mix_force(const MixArgs& M, int icp, double p, double a1,
                                          double a3, double b1, double b3, double fx[4],
                                          double fy[4]) {
  forceP(p, a1, a3, b1, b3, fx, fy);
  const double4 qx = reinterpret_cast<const double4*>(M.cpviscx)[icp];
  const double4 qy = reinterpret_cast<const double4*>(M.cpviscy)[icp];
  fx[0] = fx[0] + qx.x; fx[1] = fx[1] + qx.y; fx[2] = fx[2] + qx.z; fx[3] = fx[3] + qx.w;
  fy[0] = fy[0] + qy.x; fy[1] = fy[1] + qy.y; fy[2] = fy[2] + qy.z; fy[3] = fy[3] + qy.w;
}

// hydro_dr_getcs2 for a mixed element (hydro_dr_getcs2.f90:40-42,
// utils_kn_average12): sum over components of frmass*(cpcs2 + cpvisc)
__device__ __forceinline__ double mix_cs2(const MixArgs& M, int imx) {
  int icp = M.imxfcp[imx] - 1;
  const int n = M.imxncp[imx];
  double w1 = 0.0;
  for (int ii = 0; ii < n; ++ii, ++icp) w1 = w1 + M.frmass[icp] * (M.cpcs2[icp] + M.cpvisc[icp]);
  return w1;
}

__device__ __forceinline__ void st4c(double* a, int icp, const double v[4]) {
  reinterpret_cast<double4*>(a)[icp] = make_double4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void ld4c(const double* a, int icp, double v[4]) {
  const double4 q = reinterpret_cast<const double4*>(a)[icp];
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}

// launches of the per-driver component terms (bl_mix.hip); with ts non-null
// the kernels are gated like the ALE kernels of a captured step
int mix_geometry(Handle* h, const DevTimestep* ts);
int mix_density(Handle* h, const DevTimestep* ts);
int mix_eos(Handle* h, const DevTimestep* ts);
// geometry + density + EOS of the components in one pass (the component part
// of ale_dr_hydroupdate) from the cell geometry of the node positions x, y;
// gated by ts like the remap
int mix_hydroupdate(Handle* h, const double* x, const double* y, const DevTimestep* ts);
int mix_initmass(Handle* h);
int mix_viscinit(Handle* h);
int mix_force(Handle* h);
int mix_energy(Handle* h, double dts);
int mix_average_force(Handle* h);
// the component corner velocities cpu/cpv of the last corrector getenergy
// into rscratch23/24 (the reference's lagstep aliases cpu => rscratch23,
// cpv => rscratch24, dataAPI_id.f90:94-96), from ubar in rscratch14/15
int mix_cpu_overlay(Handle* h, const DevTimestep* ts);


// ale_dr_hydroupdate's component part for mixed cell imx, from the cell's
// iso terms g on the remapped mesh: component volume, density and EOS
// (component mass and energy are not remapped)
__device__ __forceinline__ void mix_hydroupdate_iso(const MixArgs& M, int imx, const Iso& g,
                                                    const int32_t* __restrict__ eos_itype,
                                                    const double* __restrict__ eos_param,
                                                    double pcut, double ccut) {
  const double w1 = g.a1, w2 = g.a3, w3 = g.b1, w4 = g.b3;
  int icp = M.imxfcp[imx] - 1;
  for (int ii = 0; ii < M.imxncp[imx]; ++ii, ++icp) {
    M.cpa1[icp] = w1; M.cpa3[icp] = w2; M.cpb1[icp] = w3; M.cpb3[icp] = w4;
    const double cvol = (4.0 * (w1 * w4 - w2 * w3)) * M.frvolume[icp];
    M.cpvolume[icp] = cvol;
    const double r = M.cpmass[icp] / cvol;
    M.cpdensity[icp] = r;
    int mat = M.icpmat[icp];
    mat = mat > 0 ? mat : 0;
    const int et = eos_itype[mat];
    const double* ep = eos_param + BL_NEOSPARAM * mat;
    const double e = M.cpenergy[icp];
    M.cppressure[icp] = eos_pressure(et, ep, r, e, pcut);
    M.cpcs2[icp] = eos_cs2(et, ep, r, e, pcut, ccut);
  }
}

// the same with the cell's nodes read from the remapped positions
__device__ __forceinline__ void mix_hydroupdate_cell(const MixArgs& M, int imx,
                                                     const int32_t* __restrict__ ielnd,
                                                     const double* __restrict__ ndx,
                                                     const double* __restrict__ ndy,
                                                     const int32_t* __restrict__ eos_itype,
                                                     const double* __restrict__ eos_param,
                                                     double pcut, double ccut) {
  const int iel = M.imxel[imx] - 1;
  const int4 q = reinterpret_cast<const int4*>(ielnd)[iel];
  const double x[4] = {ndx[q.x - 1], ndx[q.y - 1], ndx[q.z - 1], ndx[q.w - 1]};
  const double y[4] = {ndy[q.x - 1], ndy[q.y - 1], ndy[q.z - 1], ndy[q.w - 1]};
  Iso g;
  getiso(x, y, g);
  mix_hydroupdate_iso(M, imx, g, eos_itype, eos_param, pcut, ccut);
}

}  // namespace bl
