// bl_ale.h -- per-element pieces of the ALE remap shared by the per-driver
// kernels (bl_ale.hip) and the device-resident step (bl_fused.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bl_math.h"

namespace bl {

struct NbFc {
  int nb[4], fc[4];
};
__device__ __forceinline__ NbFc load_nbfc(int e, const int32_t* ielel, const int32_t* ielfc) {
  int4 a = reinterpret_cast<const int4*>(ielel)[e];
  int4 b = reinterpret_cast<const int4*>(ielfc)[e];
  NbFc r;
  r.nb[0] = a.x - 1; r.nb[1] = a.y - 1; r.nb[2] = a.z - 1; r.nb[3] = a.w - 1;
  r.fc[0] = b.x - 1; r.fc[1] = b.y - 1; r.fc[2] = b.z - 1; r.fc[3] = b.w - 1;
  return r;
}

// the same from the packed table of the device-resident step: per face
// (neighbour - 1) << 2 | (its face - 1), one int32 instead of two
__device__ __forceinline__ NbFc load_nbfc_packed(int e, const int32_t* nbfc) {
  const int4 a = reinterpret_cast<const int4*>(nbfc)[e];
  NbFc r;
  r.nb[0] = a.x >> 2; r.nb[1] = a.y >> 2; r.nb[2] = a.z >> 2; r.nb[3] = a.w >> 2;
  r.fc[0] = a.x & 3; r.fc[1] = a.y & 3; r.fc[2] = a.z & 3; r.fc[3] = a.w & 3;
  return r;
}

// ale_kn_sumflux for one element (ale_kn_advectors.f90:216-244)
__device__ __forceinline__ double sumflux_el(int e, int id1, int id2, const int nb[4],
                                             const int fc[4], const double* __restrict__ fl) {
  double acc = 0.0;
  for (int i1 = id1; i1 <= id2; ++i1) {
    int i2 = i1 + 2;
    int e1 = nb[i1], e2 = nb[i2];
    double w1 = fl[4 * e1 + fc[i1]];
    double w2 = fl[4 * e2 + fc[i2]];
    if (e1 == e) w1 = 0.0;
    if (e2 == e) w2 = 0.0;
    acc = acc - fl[4 * e + i1] - fl[4 * e + i2] + w1 + w2;
  }
  return acc;
}

// ale_kn_fluxelvl for one element (ale_kn_advectors.f90:30-81), with the
// element's own face values fd4 in registers; out[4] set to 0 then the two
// faces of each pass
__device__ __forceinline__ void fluxelvl_regs(int e, int id1, int id2, const int nb[4],
                                              const int fc[4], const double* __restrict__ cb,
                                              const double fd4[4],
                                              const double* __restrict__ var, double out[4]) {
  out[0] = 0.0; out[1] = 0.0; out[2] = 0.0; out[3] = 0.0;
  const double rv = var[e];
  for (int i1 = id1; i1 <= id2; ++i1) {
    int i2 = i1 + 2;
    int e2 = nb[i2], j2 = fc[i2];
    double r3 = cb[4 * e + i2] + cb[4 * e + ((i2 + 1) & 3)];
    double w5 = r3 + cb[4 * e2 + j2] + cb[4 * e2 + ((j2 + 1) & 3)];
    int e1 = nb[i1];
    j2 = fc[i1];
    double r4 = cb[4 * e + i1] + cb[4 * e + i1 + 1];
    double w6 = r4 + cb[4 * e1 + j2] + cb[4 * e1 + ((j2 + 1) & 3)];
    double f1, f2;
    fluxel_pair(r3, w5, r4, w6, rv, var[e2], var[e1], fd4[i1], fd4[i2], f1, f2);
    out[i1] = f1;
    out[i2] = f2;
  }
}

// ale_kn_sumflux for one element with its own face values and the matching
// face values of its neighbours in registers (nbv[i]: the neighbour across
// face i, ignored at a boundary face)
__device__ __forceinline__ double sumflux_regs(int e, int id1, int id2, const int nb[4],
                                               const double own[4], const double nbv[4]) {
  double acc = 0.0;
  for (int i1 = id1; i1 <= id2; ++i1) {
    int i2 = i1 + 2;
    double w1 = nbv[i1], w2 = nbv[i2];
    if (nb[i1] == e) w1 = 0.0;
    if (nb[i2] == e) w2 = 0.0;
    acc = acc - own[i1] - own[i2] + w1 + w2;
  }
  return acc;
}

// ale_kn_fluxelvl for one element (ale_kn_advectors.f90:30-81) with the
// element's face values read from fd
__device__ __forceinline__ void fluxelvl_el(int e, int id1, int id2, const int nb[4],
                                            const int fc[4], const double* __restrict__ cb,
                                            const double* __restrict__ fd,
                                            const double* __restrict__ var, double out[4]) {
  const double fd4[4] = {fd[4 * e], fd[4 * e + 1], fd[4 * e + 2], fd[4 * e + 3]};
  fluxelvl_regs(e, id1, id2, nb, fc, cb, fd4, var, out);
}

// ale_kn_fluxndvl for one element (ale_kn_advectors.f90:83-153): corner
// fluxes of a nodal quantity var (corner array) with the limited gradient
// across the corners of the element and its face neighbours; cb is the corner
// basis, cdb the corner basis flux
__device__ __forceinline__ void fluxndvl_el(int e, const NbFc& c, const double* __restrict__ cb,
                                            const double* __restrict__ cdb,
                                            const double* __restrict__ var, double out[4]) {
  out[0] = 0.0; out[1] = 0.0; out[2] = 0.0; out[3] = 0.0;
#pragma unroll
  for (int l = 0; l < 2; ++l) {        // ilfcl-1
    const int lr = l + 2;              // ilfcr-1
#pragma unroll
    for (int icn = 1; icn <= 2; ++icn) {
      const int ilndl = l + icn - 1;               // 0-based ilfcl+icn-1
      const int ilndr = (lr + 1 - icn + 1) & 3;    // MOD(ilfcr-icn+1,4)+1, 0-based
      double rd = 0.0;
      const int iell = c.nb[l], ielr = c.nb[lr];
      const int ifcl = c.fc[l] + 1, ifcr = c.fc[lr] + 1;  // 1-based faces
      const int ii = l + 2 * (icn - 1);
      const double cd = cdb[4 * e + ii];
      if (cd > 0.0) {
        int ilnndl = (ifcl + icn) % 4;                // MOD(ifcl+icn,4)+1, 0-based
        int ilnndr = (ifcl - icn + 1) % 4;            // MOD(ifcl-icn+1,4)+1, 0-based
        double rv = var[4 * e + ilndl];
        rd = cb[4 * e + ilndl] - 0.5 * cd;
        double w5 = cb[4 * e + ilndl] + cb[4 * e + ilndr];
        double w6 = cb[4 * iell + ilnndl] + cb[4 * iell + ilnndr];
        double w1 = rv - var[4 * iell + ilnndl];
        double w2 = var[4 * e + ilndr] - rv;
        double w3 = fabs(w1), w4 = fabs(w2);
        double w7 = dsign(1.0, w2);
        double w8 = (w2 * w6 * w6 + w1 * w5 * w5) / (w5 * w6 * (w5 + w6));
        double grad = w7 * dmin(dmin(fabs(w8), w3 / w6), w4 / w5);
        if (w1 * w2 <= 0.0) grad = 0.0;
        rd = cd * (rv + grad * rd);
      }
      if (cd < 0.0) {
        int ilnndl = (ifcr + icn - 2) % 4;                      // MOD(ifcr+icn-2,4)+1
        int ilnndr = (((ifcr - icn - 1) % 4) + 4) % 4;          // MODULO(ifcr-icn-1,4)+1
        double rv = var[4 * e + ilndr];
        rd = cb[4 * e + ilndr] + 0.5 * cd;
        double w5 = cb[4 * e + ilndl] + cb[4 * e + ilndr];
        double w6 = cb[4 * ielr + ilnndl] + cb[4 * ielr + ilnndr];
        double w1 = rv - var[4 * e + ilndl];
        double w2 = var[4 * ielr + ilnndr] - rv;
        double w3 = fabs(w1), w4 = fabs(w2);
        double w7 = dsign(1.0, w2);
        double w8 = (w1 * w6 * w6 + w2 * w5 * w5) / (w5 * w6 * (w5 + w6));
        double grad = -(w7 * dmin(dmin(fabs(w8), w3 / w5), w4 / w6));
        if (w1 * w2 <= 0.0) grad = 0.0;
        rd = cd * (rv + grad * rd);
      }
      out[ilndl] = out[ilndl] - rd;
      out[ilndr] = out[ilndr] + rd;
    }
  }
}

// The two limited upwind terms of fluxndvl_el for one component: the flux
// through a left face (cd > 0) and through a right face (cd < 0). rv is the
// upwind corner's value, vnb the neighbour's corner beyond the face, vown
// the element's other corner on the face, rd the corner-basis term.
__device__ __forceinline__ double fluxnd_left(double cd, double rd, double w5, double w6,
                                              double rv, double vnb, double vown) {
  const double w1 = rv - vnb;
  const double w2 = vown - rv;
  const double w3 = fabs(w1), w4 = fabs(w2);
  const double w7 = dsign(1.0, w2);
  const double w8 = (w2 * w6 * w6 + w1 * w5 * w5) / (w5 * w6 * (w5 + w6));
  double grad = w7 * dmin(dmin(fabs(w8), w3 / w6), w4 / w5);
  if (w1 * w2 <= 0.0) grad = 0.0;
  return cd * (rv + grad * rd);
}
__device__ __forceinline__ double fluxnd_right(double cd, double rd, double w5, double w6,
                                               double rv, double vown, double vnb) {
  const double w1 = rv - vown;
  const double w2 = vnb - rv;
  const double w3 = fabs(w1), w4 = fabs(w2);
  const double w7 = dsign(1.0, w2);
  const double w8 = (w1 * w6 * w6 + w2 * w5 * w5) / (w5 * w6 * (w5 + w6));
  double grad = -(w7 * dmin(dmin(fabs(w8), w3 / w5), w4 / w6));
  if (w1 * w2 <= 0.0) grad = 0.0;
  return cd * (rv + grad * rd);
}

// fluxndvl_el for u and v together, the corner values read from the nodes:
// the corner value of element k at corner j is the velocity of node
// ielnd(j, k) (1-based), which is what a corner gather of u and v before the
// update would hold. The basis terms are shared by the two components; each
// component's operations are fluxndvl_el's, in its order.
__device__ __forceinline__ void fluxndvl_uv(int e, const NbFc& c,
                                            const int32_t* __restrict__ ielnd,
                                            const double* __restrict__ cb,
                                            const double* __restrict__ cdb,
                                            const double* __restrict__ u,
                                            const double* __restrict__ v, double ou[4],
                                            double ov[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) { ou[j] = 0.0; ov[j] = 0.0; }
#pragma unroll
  for (int l = 0; l < 2; ++l) {
    const int lr = l + 2;
#pragma unroll
    for (int icn = 1; icn <= 2; ++icn) {
      const int ilndl = l + icn - 1;
      const int ilndr = (lr + 1 - icn + 1) & 3;
      double du = 0.0, dv = 0.0;
      const int ii = l + 2 * (icn - 1);
      const double cd = cdb[4 * e + ii];
      if (cd > 0.0) {
        const int iell = c.nb[l], ifcl = c.fc[l] + 1;
        const int ilnndl = (ifcl + icn) % 4;
        const int ilnndr = (ifcl - icn + 1) % 4;
        const int na = ielnd[4 * e + ilndl] - 1, nb = ielnd[4 * e + ilndr] - 1;
        const int nn = ielnd[4 * iell + ilnndl] - 1;
        const double rd = cb[4 * e + ilndl] - 0.5 * cd;
        const double w5 = cb[4 * e + ilndl] + cb[4 * e + ilndr];
        const double w6 = cb[4 * iell + ilnndl] + cb[4 * iell + ilnndr];
        du = fluxnd_left(cd, rd, w5, w6, u[na], u[nn], u[nb]);
        dv = fluxnd_left(cd, rd, w5, w6, v[na], v[nn], v[nb]);
      }
      if (cd < 0.0) {
        const int ielr = c.nb[lr], ifcr = c.fc[lr] + 1;
        const int ilnndl = (ifcr + icn - 2) % 4;
        const int ilnndr = (((ifcr - icn - 1) % 4) + 4) % 4;
        const int na = ielnd[4 * e + ilndr] - 1, nb = ielnd[4 * e + ilndl] - 1;
        const int nn = ielnd[4 * ielr + ilnndr] - 1;
        const double rd = cb[4 * e + ilndr] + 0.5 * cd;
        const double w5 = cb[4 * e + ilndl] + cb[4 * e + ilndr];
        const double w6 = cb[4 * ielr + ilnndl] + cb[4 * ielr + ilnndr];
        du = fluxnd_right(cd, rd, w5, w6, u[na], u[nb], u[nn]);
        dv = fluxnd_right(cd, rd, w5, w6, v[na], v[nb], v[nn]);
      }
      ou[ilndl] = ou[ilndl] - du;
      ou[ilndr] = ou[ilndr] + du;
      ov[ilndl] = ov[ilndl] - dv;
      ov[ilndr] = ov[ilndr] + dv;
    }
  }
}

}  // namespace bl
