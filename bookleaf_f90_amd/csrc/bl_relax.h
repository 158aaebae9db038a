// bl_relax.h -- one node of one sweep of the node relaxation, shared by the
// relaxation kernels and the learned relaxation's features.
#pragma once

#include <hip/hip_runtime.h>

#include "bl_internal.h"

namespace bl {

struct RelaxArgs {
  int nnd, type;
  double om, tol, zerocut;
  const int32_t* ring;
  const int32_t* only_if;  // non-null: the sweeps run only when *only_if != 0
};

// one node of one sweep (setup_mesh.F90:797-863); (xo, yo) the position
// after it.  rr of the generator is y, ss is x.
__device__ __forceinline__ void relax_node(const RelaxArgs& R, int n, const double* __restrict__ xi,
                                           const double* __restrict__ yi, double& xo,
                                           double& yo) {
  const double xn = xi[n], yn = yi[n];
  xo = xn;
  yo = yn;
  const int4 qa = reinterpret_cast<const int4*>(R.ring)[2 * n];
  const int4 qb = reinterpret_cast<const int4*>(R.ring)[2 * n + 1];
  if (qa.x < 0) return;
  const int nb[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
  double x[8], y[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    x[i] = xi[nb[i]];
    y[i] = yi[nb[i]];
  }
  double r1, r2, r3, r4, r5, r6, r7, r8;
  if (R.type == BL_RELAX_EQUI) {
    // (ll,kp) = 4, (ll,km) = 3, (lp,kk) = 1, (lm,kk) = 6
    const double r_psi = 0.5 * (y[4] - y[3]);
    const double s_psi = 0.5 * (x[4] - x[3]);
    const double r_phi = 0.5 * (y[1] - y[6]);
    const double s_phi = 0.5 * (x[1] - x[6]);
    const double zeta = s_psi * r_phi - s_phi * r_psi;
    if (fabs(zeta) < R.zerocut) return;
    const double alph = s_psi * s_psi + r_psi * r_psi;
    const double beta = s_phi * s_psi + r_phi * r_psi;
    const double gamm = s_phi * s_phi + r_phi * r_phi;
    r1 = 0.5 * beta; r2 = alph; r3 = -r1; r4 = gamm;
    r5 = r4; r6 = r3; r7 = r2; r8 = r1;
    double fac = r1 + r2 + r3 + r4 + r5 + r6 + r7 + r8;
    if (fac < 1.0e-24) {
      fac = fac * 1.0e14;
      r1 = r1 * 1.0e14; r2 = r2 * 1.0e14; r3 = r3 * 1.0e14; r4 = r4 * 1.0e14;
      r5 = r5 * 1.0e14; r6 = r6 * 1.0e14; r7 = r7 * 1.0e14; r8 = r8 * 1.0e14;
    }
    r1 = r1 / fac; r2 = r2 / fac; r3 = r3 / fac; r4 = r4 / fac;
    r5 = r5 / fac; r6 = r6 / fac; r7 = r7 / fac; r8 = r8 / fac;
  } else {
    // LIN1: r_wgt(2) = r_wgt(7) = 1; LIN2: r_wgt(4) = r_wgt(5) = 1; each
    // divided by their sum (setup_mesh.F90:626-660)
    const double w = 1.0 / 2.0;
    r1 = 0.0; r2 = 0.0; r3 = 0.0; r4 = 0.0; r5 = 0.0; r6 = 0.0; r7 = 0.0; r8 = 0.0;
    if (R.type == BL_RELAX_LIN1) { r2 = w; r7 = w; } else { r4 = w; r5 = w; }
  }
  // s_i = r_i for every type here
  const double dr = r1 * y[0] + r2 * y[1] + r3 * y[2] + r4 * y[3] + r5 * y[4] + r6 * y[5] +
                    r7 * y[6] + r8 * y[7] - yn;
  const double ds = r1 * x[0] + r2 * x[1] + r3 * x[2] + r4 * x[3] + r5 * x[4] + r6 * x[5] +
                    r7 * x[6] + r8 * x[7] - xn;
  if ((fabs(dr) > R.tol) || (fabs(ds) > R.tol)) {
    yo = yn + R.om * dr;
    xo = xn + R.om * ds;
  }
}

// the arguments of the relaxation set on the handle (bl_set_relaxation)
RelaxArgs relax_args(Handle* h);

}  // namespace bl
