// bl_math.h -- per-element / per-node FP64 arithmetic of BookLeaf's hot path
// for gfx950, shared by the per-driver kernels and the fused step kernels.
//
// Every function keeps the reference's left-to-right association and uses
// only IEEE-exact operations (+ - * / sqrt, no contraction: the library is
// built with -ffp-contract=off), so results are bit-identical to the
// reference built without FMA.  Corner/face loops are fully unrolled; arrays
// of 4 stay in registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bl {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup b runs on XCD b mod 8) and each XCD has its own L2.  Remap so
// XCD x processes one contiguous 1/8 of the index range: mesh neighbours
// (the element rows shared by adjacent node rows, the neighbour elements of
// the limiter and the remap) then hit the same L2 instead of being fetched
// from HBM once per XCD.  A bijection on [0, gridDim.x).
constexpr int kNumXcd = 8;
// (xcd_block_of(b, n): the same over the first n workgroups of a grid that
// appends workgroups of another kind after them)
__device__ __forceinline__ int xcd_block_of(int b, int n) {
  const int x = b % kNumXcd, local = b / kNumXcd;
  const int q = n / kNumXcd, r = n % kNumXcd;
  return x * q + (x < r ? x : r) + local;
}
__device__ __forceinline__ int xcd_block() { return xcd_block_of(blockIdx.x, gridDim.x); }
__device__ __forceinline__ int xcd_index() { return xcd_block() * blockDim.x + threadIdx.x; }

__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
// Fortran SIGN(a,b)
__device__ __forceinline__ double dsign(double a, double b) { return copysign(fabs(a), b); }

// geometry_kn_getiso + geometry_kn_getvolume (geometry_kn_volume.f90:30-86)
struct Iso {
  double a1, a2, a3, b1, b2, b3, wt[4], vol;
};
__device__ __forceinline__ void getiso(const double x[4], const double y[4], Iso& g) {
  const double ONEBYNINE = 1.0 / 9.0;
  g.a1 = 0.25 * (-x[0] + x[1] + x[2] - x[3]);
  g.a2 = 0.25 * (x[0] - x[1] + x[2] - x[3]);
  g.a3 = 0.25 * (-x[0] - x[1] + x[2] + x[3]);
  g.b1 = 0.25 * (-y[0] + y[1] + y[2] - y[3]);
  g.b2 = 0.25 * (y[0] - y[1] + y[2] - y[3]);
  g.b3 = 0.25 * (-y[0] - y[1] + y[2] + y[3]);
  g.wt[0] = ONEBYNINE * ((3.0 * g.b3 - g.b2) * (3.0 * g.a1 - g.a2) -
                         (3.0 * g.a3 - g.a2) * (3.0 * g.b1 - g.b2));
  g.wt[1] = ONEBYNINE * ((3.0 * g.b3 + g.b2) * (3.0 * g.a1 - g.a2) -
                         (3.0 * g.a3 + g.a2) * (3.0 * g.b1 - g.b2));
  g.wt[2] = ONEBYNINE * ((3.0 * g.b3 + g.b2) * (3.0 * g.a1 + g.a2) -
                         (3.0 * g.a3 + g.a2) * (3.0 * g.b1 + g.b2));
  g.wt[3] = ONEBYNINE * ((3.0 * g.b3 - g.b2) * (3.0 * g.a1 + g.a2) -
                         (3.0 * g.a3 - g.a2) * (3.0 * g.b1 + g.b2));
  g.vol = 4.0 * (g.a1 * g.b3 - g.a3 * g.b1);
}

// geometry_kn_length.f90:30-139 (dlm / dln and helpers)
__device__ __forceinline__ double g_denom(double x1, double y1, double x2, double y2) {
  double w1 = y1 - y2, w2 = x1 - x2;
  return w1 * w1 + w2 * w2;
}
__device__ __forceinline__ double g_distpp(double x3, double y3, double x4, double y4,
                                           double x1, double y1) {
  double w1 = 0.5 * (x3 + x4) - x1, w2 = 0.5 * (y3 + y4) - y1;
  return w1 * w1 + w2 * w2;
}
__device__ __forceinline__ double g_distpl(double x3, double y3, double x4, double y4,
                                           double x1, double y1, double x2, double y2) {
  double r = 0.5 * (y1 - y2) * (x3 + x4) + 0.5 * (y3 + y4) * (x2 - x1) + y2 * x1 - y1 * x2;
  return r * r;
}
__device__ __forceinline__ double dln_side(const double* x, const double* y, int a, int b,
                                           int c, int d, double zcut) {
  double w1 = g_denom(x[c], y[c], x[d], y[d]);
  return (w1 < zcut) ? g_distpp(x[a], y[a], x[b], y[b], x[c], y[c])
                     : g_distpl(x[a], y[a], x[b], y[b], x[c], y[c], x[d], y[d]) / w1;
}
// MINVAL of geometry_kn_dln or geometry_kn_dlm
__device__ __forceinline__ double minlength(const double x[4], const double y[4], double zcut,
                                            bool midlength) {
  if (midlength) {
    double x1 = x[0] + x[1], x2 = x[2] + x[3], y1 = y[0] + y[1], y2 = y[2] + y[3];
    x1 = 0.5 * (x1 - x2);
    y1 = 0.5 * (y1 - y2);
    double r0 = x1 * x1 + y1 * y1;
    x1 = x[2] + x[1]; x2 = x[0] + x[3]; y1 = y[2] + y[1]; y2 = y[0] + y[3];
    x1 = 0.5 * (x1 - x2);
    y1 = 0.5 * (y1 - y2);
    double r1 = x1 * x1 + y1 * y1;
    return dmin(dmin(dmin(r0, r1), r0), r1);
  }
  double r0 = dln_side(x, y, 0, 1, 2, 3, zcut);
  double r1 = dln_side(x, y, 1, 2, 3, 0, zcut);
  double r2 = dln_side(x, y, 2, 3, 0, 1, zcut);
  double r3 = dln_side(x, y, 3, 0, 1, 2, zcut);
  return dmin(dmin(dmin(r0, r1), r2), r3);
}

// eos_kn_getpressure / eos_kn_getcs2 (eos_kn_forms.f90:28-106); param points
// at the 6 parameters of material imat = max(mat,0)
// (p0: p[0], which a caller may have loaded ahead)
__device__ __forceinline__ double eos_pressure(int type, double p0, const double* p,
                                               double density, double energy, double pcut) {
  double pressure, t1, t2, t3, t4, t5;
  switch (type) {
    case 0: pressure = p0; break;
    case 1: pressure = energy * density * (p0 - 1.0); break;
    case 2:
      t1 = density / p[2];
      pressure = p[0] * (pow(t1, p[1]) - 1.0);
      pressure = dmax(pressure, p[3]);
      break;
    case 3:
      t1 = p[3] * p[5] / density;
      t2 = p[4] * p[5] / density;
      t3 = p[0] * density * energy;
      t4 = (1.0 - p[0] / t1) * p[1] * exp(-t1);
      t5 = (1.0 - p[0] / t2) * p[2] * exp(-t2);
      pressure = t3 + t4 + t5;
      break;
    default: pressure = -1.0;
  }
  if (fabs(pressure) < pcut) pressure = 0.0;
  return pressure;
}
__device__ __forceinline__ double eos_pressure(int type, const double* p, double density,
                                               double energy, double pcut) {
  return eos_pressure(type, p[0], p, density, energy, pcut);
}
__device__ __forceinline__ double eos_cs2(int type, const double* p, double density,
                                          double energy, double pcut, double ccut) {
  double cs2, t1, t2, t3, t4, t5;
  switch (type) {
    case 0: cs2 = ccut; break;
    case 1: cs2 = p[0] * (p[0] - 1.0) * energy; break;
    case 2:
      t1 = density / p[2];
      t2 = p[1] - 1.0;
      cs2 = (p[0] * p[1]) / p[2];
      cs2 = cs2 * pow(t1, t2);
      break;
    case 3:
      t1 = p[5] / density;
      t2 = eos_pressure(3, p, density, energy, pcut);
      t3 = p[3] * t1;
      t4 = p[0] / p[3] + p[0] * t1 - t3 * t1;
      t4 = t4 * p[1] * exp(-t3);
      t3 = p[4] * t1;
      t5 = p[0] / p[4] + p[0] * t1 - t3 * t1;
      t5 = t5 * p[2] * exp(-t3);
      cs2 = p[0] * t2 / density + p[0] * energy - t4 - t5;
      break;
    default: cs2 = ccut;
  }
  return dmax(cs2, ccut);
}

// One limiter half-pass of hydro_kn_limiteviscosity for edge pair (is1 or is2)
// (hydro_kn_edgeviscosity.f90:103-145): returns (w2*w1, w3*w1) for the two
// neighbour ratios.  Inputs: own edge du,dv,dx,dy and the two neighbour
// edges' du,dv,dx,dy.
__device__ __forceinline__ void limiter_ratios(double du, double dv, double dx, double dy,
                                               double n1u, double n1v, double n1x, double n1y,
                                               double n2u, double n2v, double n2x, double n2y,
                                               double zerocut, double& r1, double& r2) {
  double w1 = du, w2 = dv, w3 = dx, w4 = dy;
  double den = sqrt(w1 * w1 + w2 * w2);
  den = 1.0 / dmax(den, zerocut);
  double uhat = w1 * den, vhat = w2 * den;
  den = sqrt(w3 * w3 + w4 * w4);
  den = 1.0 / dmax(den, zerocut);
  double xhat = w3 * den, yhat = w4 * den;
  den = w3 * xhat + w4 * yhat;
  w1 = (w1 * uhat + w2 * vhat) / dsign(dmax(fabs(den), zerocut), den);
  w1 = 1.0 / dsign(dmax(fabs(w1), zerocut), w1);
  den = n1x * xhat + n1y * yhat;
  w2 = (n1u * uhat + n1v * vhat) / dsign(dmax(fabs(den), zerocut), den);
  r1 = w2 * w1;
  den = n2x * xhat + n2y * yhat;
  w3 = (n2u * uhat + n2v * vhat) / dsign(dmax(fabs(den), zerocut), den);
  r2 = w3 * w1;
}

// Edge coefficient of the "apply limiter" loop (hydro_kn_edgeviscosity.f90:188-214)
__device__ __forceinline__ double limiter_edge(double s1, double s2, double du, double dv,
                                               double cs2_i, double cs2_n, double rho_i,
                                               double rho_n, double cvisc1, double cvisc2,
                                               double zerocut, double& elvisc) {
  double w1 = cvisc1 * sqrt(0.5 * (cs2_i + cs2_n));
  double w2 = dmin(dmin(dmin(0.5 * (s1 + s2), 2.0 * s1), 2.0 * s2), 1.0);
  w2 = dmax(0.0, w2);
  double w4 = sqrt(du * du + dv * dv);
  double w3 = 0.5 * (1.0 - w2) * (rho_i + rho_n) * (w1 + cvisc2 * w4);
  w4 = 2.0 * w4 / dmax(rho_i, zerocut);
  elvisc = dmax(elvisc, w3 * w4);
  return w3;
}

// hydro_kn_geteviscosity for one element (hydro_kn_edgeviscosity.f90:219-282):
// edge coefficients in ed[4] (edviscx = edviscy on entry), out: corner forces
__device__ __forceinline__ void geteviscosity(const double x[4], const double y[4],
                                              const double u[4], const double v[4],
                                              const double edx_in[4], const double edy_in[4],
                                              double zerocut, double qx[4], double qy[4]) {
  double edx[4], edy[4];
  double xc = 0.25 * (x[0] + x[1] + x[2] + x[3]);
  double yc = 0.25 * (y[0] + y[1] + y[2] + y[3]);
#pragma unroll
  for (int is = 0; is < 4; ++is) {
    int in = (is + 1) & 3;
    double w1 = x[is], w2 = x[in];
    double w3 = 0.5 * (w1 + w2);
    w1 = w2 - w1;
    w2 = xc;
    double w4 = y[is], w5 = y[in];
    double w6 = 0.5 * (w4 + w5);
    w4 = w5 - w4;
    w5 = yc;
    double w7 = sqrt((w2 - w3) * (w2 - w3) + (w5 - w6) * (w5 - w6));
    double w8 = sqrt(w1 * w1 + w4 * w4);
    double den = 1.0 / w7;
    double xhat = (w5 - w6) * den;
    double yhat = (w3 - w2) * den;
    den = 1.0 / w8;
    w1 = w1 * den;
    w2 = w4 * den;
    w3 = xhat * w1 + yhat * w2;
    den = -(dsign(1.0, w3) * w7);
    xhat = xhat * den;
    yhat = yhat * den;
    double uhat = u[in] - u[is];
    double vhat = v[in] - v[is];
    w5 = sqrt((uhat * uhat) + (vhat * vhat));
    w6 = uhat * xhat + vhat * yhat;
    den = w6 / dmax(w5, zerocut);
    edx[is] = edx_in[is] * uhat * den;
    edy[is] = edy_in[is] * vhat * den;
    if (w5 <= zerocut || w6 <= zerocut || w7 <= zerocut || w8 <= zerocut) {
      edx[is] = 0.0;
      edy[is] = 0.0;
    }
  }
  qx[0] = edx[0] - edx[3]; qx[1] = edx[1] - edx[0];
  qx[2] = edx[2] - edx[1]; qx[3] = edx[3] - edx[2];
  qy[0] = edy[0] - edy[3]; qy[1] = edy[1] - edy[0];
  qy[2] = edy[2] - edy[1]; qy[3] = edy[3] - edy[2];
}

// hydro_kn_getdtdiv per-element term (hydro_kn_getdt.f90:106-113)
__device__ __forceinline__ double divterm(const double u[4], const double v[4], double a1,
                                          double a3, double b1, double b3, double vol) {
  double w1 = u[0] * (-b3 + b1) + v[0] * (a3 - a1) + u[1] * (b3 + b1) + v[1] * (-a3 - a1) +
              u[2] * (b3 - b1) + v[2] * (-a3 + a1) + u[3] * (-b3 - b1) + v[3] * (a3 + a1);
  return fabs(w1) / vol;
}

// hydro_kn_getforceP (hydro_kn_getforce.f90:39-50)
__device__ __forceinline__ void forceP(double p, double a1, double a3, double b1, double b3,
                                       double fx[4], double fy[4]) {
  fx[0] = p * (-b3 + b1); fx[1] = p * (b3 + b1);
  fx[2] = p * (b3 - b1);  fx[3] = p * (-b3 - b1);
  fy[0] = p * (a3 - a1);  fy[1] = p * (-a3 - a1);
  fy[2] = p * (-a3 + a1); fy[3] = p * (a3 + a1);
}

// hydro_kn_getforceHG per element (hydro_kn_getforce.f90:86-104); w4 = 1/dt
__device__ __forceinline__ void forceHG(double kappa, double rho, double area, double w4,
                                        const double u[4], const double v[4], double fx[4],
                                        double fy[4]) {
  double w2 = u[0] - u[1] + u[2] - u[3];
  double w3 = v[0] - v[1] + v[2] - v[3];
  double w1 = -(kappa * rho * area * w4);
  w2 = w1 * w2;
  w3 = w1 * w3;
  fx[0] = fx[0] + w2; fx[1] = fx[1] - w2; fx[2] = fx[2] + w2; fx[3] = fx[3] - w2;
  fy[0] = fy[0] + w3; fy[1] = fy[1] - w3; fy[2] = fy[2] + w3; fy[3] = fy[3] - w3;
}

// hydro_kn_getforceSP per element (hydro_kn_getforce.f90:123-186)
__device__ __forceinline__ void forceSP(double w1, double rho, double cs2, const double x[4],
                                        const double y[4], const double spm[4], double fx[4],
                                        double fy[4]) {
  double x3 = 0.25 * (x[0] + x[1] + x[2] + x[3]);
  double y3 = 0.25 * (y[0] + y[1] + y[2] + y[3]);
  double lfx[4][4], lfy[4][4];
#pragma unroll
  for (int j1 = 0; j1 < 4; ++j1) {
    double x1 = x[j1], y1 = y[j1];
    int j2 = (j1 + 1) & 3;
    double x2 = 0.5 * (x1 + x[j2]);
    double y2 = 0.5 * (y1 + y[j2]);
    j2 = (j1 + 3) & 3;
    double x4 = 0.5 * (x1 + x[j2]);
    double y4 = 0.5 * (y1 + y[j2]);
    double w3 = 0.25 * (-x1 + x2 + x3 - x4);
    double w4 = 0.25 * (-x1 - x2 + x3 + x4);
    double w5 = 0.25 * (-y1 + y2 + y3 - y4);
    double w6 = 0.25 * (-y1 - y2 + y3 + y4);
    double w2 = 4.0 * (w3 * w6 - w4 * w5);
    w2 = spm[j1] / w2;
    w2 = w2 - rho;
    w2 = cs2 * w2;
    lfx[j1][0] = w2 * (w5 - w6);  lfx[j1][1] = w2 * (w5 + w6);
    lfx[j1][2] = w2 * (-w5 + w6); lfx[j1][3] = w2 * (-w5 - w6);
    lfy[j1][0] = w2 * (-w3 + w4); lfy[j1][1] = w2 * (-w3 - w4);
    lfy[j1][2] = w2 * (w3 - w4);  lfy[j1][3] = w2 * (w3 + w4);
  }
  double w2 = 0.5 * (lfx[0][3] + lfx[3][1]);
  double w3 = 0.5 * (lfx[0][1] + lfx[1][3]);
  double w4 = 0.5 * (lfx[1][1] + lfx[2][3]);
  double w5 = 0.5 * (lfx[3][3] + lfx[2][1]);
  double w6 = 0.25 * (lfx[0][2] + lfx[1][2] + lfx[2][2] + lfx[3][2]);
  fx[0] = fx[0] + w1 * (lfx[0][0] + w2 + w3 + w6);
  fx[1] = fx[1] + w1 * (lfx[1][0] + w4 + w3 + w6);
  fx[2] = fx[2] + w1 * (lfx[2][0] + w4 + w5 + w6);
  fx[3] = fx[3] + w1 * (lfx[3][0] + w2 + w5 + w6);
  w2 = 0.5 * (lfy[0][3] + lfy[3][1]);
  w3 = 0.5 * (lfy[0][1] + lfy[1][3]);
  w4 = 0.5 * (lfy[1][1] + lfy[2][3]);
  w5 = 0.5 * (lfy[3][3] + lfy[2][1]);
  w6 = 0.25 * (lfy[0][2] + lfy[1][2] + lfy[2][2] + lfy[3][2]);
  fy[0] = fy[0] + w1 * (lfy[0][0] + w2 + w3 + w6);
  fy[1] = fy[1] + w1 * (lfy[1][0] + w4 + w3 + w6);
  fy[2] = fy[2] + w1 * (lfy[2][0] + w4 + w5 + w6);
  fy[3] = fy[3] + w1 * (lfy[3][0] + w2 + w5 + w6);
}

// hydro_kn_spmassinit per element (hydro_kn_init.f90:56-76)
__device__ __forceinline__ void spmassinit(double rho, const double x[4], const double y[4],
                                           double spm[4]) {
  double x3 = 0.25 * (x[0] + x[1] + x[2] + x[3]);
  double y3 = 0.25 * (y[0] + y[1] + y[2] + y[3]);
#pragma unroll
  for (int j1 = 0; j1 < 4; ++j1) {
    double x1 = x[j1], y1 = y[j1];
    int j2 = (j1 + 1) & 3;
    double x2 = 0.5 * (x1 + x[j2]);
    double y2 = 0.5 * (y1 + y[j2]);
    j2 = (j1 + 3) & 3;
    double x4 = 0.5 * (x1 + x[j2]);
    double y4 = 0.5 * (y1 + y[j2]);
    double w1 = 0.25 * (-x1 + x2 + x3 - x4);
    double w2 = 0.25 * (-x1 - x2 + x3 + x4);
    double w3 = 0.25 * (-y1 + y2 + y3 - y4);
    double w4 = 0.25 * (-y1 - y2 + y3 + y4);
    spm[j1] = 4.0 * rho * (w1 * w4 - w2 * w3);
  }
}

// hydro_kn_getenergy per element (hydro_kn_getenergy.f90:42-49)
__device__ __forceinline__ double energy_update(double e, double m, const double fx[4],
                                                const double fy[4], const double u[4],
                                                const double v[4], double dt, double zerocut) {
  double w1 = fx[0] * u[0] + fy[0] * v[0] + fx[1] * u[1] + fy[1] * v[1] + fx[2] * u[2] +
              fy[2] * v[2] + fx[3] * u[3] + fy[3] * v[3];
  w1 = -(w1 / dmax(m, zerocut));
  return e + w1 * dt;
}

// geometry_kn_getfluxvolume for one face (geometry_kn_volume.f90:121-136)
__device__ __forceinline__ double fluxvolume(double x1, double y1, double x2, double y2,
                                             double x3, double y3, double x4, double y4,
                                             double cut) {
  double a1 = (-x1 - x4) + (x3 + x2);
  double a3 = (-x1 + x4) + (x3 - x2);
  double b1 = (-y1 - y4) + (y3 + y2);
  double b3 = (-y1 + y4) + (y3 - y2);
  double v = 0.25 * (a1 * b3 - a3 * b1);
  return (v < cut) ? 0.0 : v;
}

// limited donor flux of ale_kn_fluxelvl for one (i1,i2) face pair
// (ale_kn_advectors.f90:54-77)
__device__ __forceinline__ void fluxel_pair(double r3, double w5, double r4, double w6,
                                            double rv, double var2, double var1, double r1,
                                            double r2, double& f1, double& f2) {
  double w1 = rv - var2;
  double w2 = var1 - rv;
  double w3 = fabs(w1), w4 = fabs(w2);
  double w7 = dsign(1.0, w2);
  double w8 = (w1 * w6 * w6 + w2 * w5 * w5) / (w5 * w6 * (w5 + w6));
  double grad = w7 * dmin(dmin(fabs(w8), w3 / w5), w4 / w6);
  if (w1 * w2 <= 0.0) grad = 0.0;
  f1 = r1 * (rv + grad * (r3 - 0.5 * r1));
  f2 = r2 * (rv - grad * (r4 - 0.5 * r2));
}

}  // namespace bl
