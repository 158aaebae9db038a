// bl_ale.hip -- ALE step drivers (src/packages/ale/driver/*.f90) on gfx950.
//
// Element kernels read neighbours through ielel/ielfc; the node sums of the
// nodal remap are deterministic CSR gathers in the reference's (element,
// corner) order (ale_kn_advect.f90:80-126,190-219, ale_kn_advectors.f90:
// 194-205), so results are bit-identical to the serial scatters.  The alestep
// aliases of dataAPI_id.f90:102-111 are kept: every kernel writes exactly the
// scratch entries the reference writes (the split advection reads entries of
// rwork1/rwork2 that the current pass does not write, so nothing else may be
// touched).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "bl_internal.h"
#include "bl_math.h"
#include "bl_mix.h"
#include "bl_ale.h"

namespace bl {

// Inside a captured step (ts != null) the kernels do nothing when the step is
// gated off or ALE is inactive (ale_active, ale_dr_utils.f90:35-51), and take
// dt from the device timestep.
#define ALE_GATE(ts) \
  if ((ts) && ((ts)->skip || !(ts)->ale_on)) return

// ale_kn_getmeshstatus (Eulerian, or node relaxation: indstatus = 2)
__global__ void k_meshstatus(int nnd, int32_t* __restrict__ st, const DevTimestep* ts) {
  ALE_GATE(ts);
  int i = xcd_index();
  if (i < nnd) st[i] = 2;
}

// ale_kn_getmeshvelocity + utils_kn_copy + geometry_kn_getvertex
// (ale_dr_getfluxvolume.f90:49-59)
__global__ void k_fluxvol_nodes(int nnd, int zeul, double dt, double* __restrict__ st4,
                                double* __restrict__ st5, double* __restrict__ st2,
                                double* __restrict__ st3, double* __restrict__ ndx,
                                double* __restrict__ ndy, const DevTimestep* ts) {
  ALE_GATE(ts);
  if (ts) dt = ts->dt;
  int i = xcd_index();
  if (i >= nnd) return;
  double u = st4[i], v = st5[i];
  if (zeul) { u = -u; v = -v; st4[i] = u; st5[i] = v; }
  double x = ndx[i], y = ndy[i];
  st2[i] = x;
  st3[i] = y;
  ndx[i] = x + dt * u;
  ndy[i] = y + dt * v;
}

// geometry_kn_getfluxvolume (geometry_kn_volume.f90:101-140)
__global__ void k_fluxvol_faces(int nel, double cut, const int32_t* __restrict__ ielnd,
                                const double* __restrict__ x0, const double* __restrict__ y0,
                                const double* __restrict__ x1, const double* __restrict__ y1,
                                double* __restrict__ fcdv, const DevTimestep* ts) {
  ALE_GATE(ts);
  int e = xcd_index();
  if (e >= nel) return;
  int4 nd4 = reinterpret_cast<const int4*>(ielnd)[e];
  int nd[4] = {nd4.x - 1, nd4.y - 1, nd4.z - 1, nd4.w - 1};
  double ax[4], ay[4], bx[4], by[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) { ax[j] = x0[nd[j]]; ay[j] = y0[nd[j]]; bx[j] = x1[nd[j]]; by[j] = y1[nd[j]]; }
  double f[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int jp = (j + 1) & 3;
    f[j] = fluxvolume(ax[j], ay[j], ax[jp], ay[jp], bx[jp], by[jp], bx[j], by[j], cut);
  }
  reinterpret_cast<double4*>(fcdv)[e] = make_double4(f[0], f[1], f[2], f[3]);
}

// advectel part 1: ale_kn_sumflux(fcdv)->rwork1, ale_kn_fluxelvl(cnwt,fcdv,rho)->fcdm
__global__ void k_el1(int nel, int id1, int id2, const int32_t* __restrict__ ielel,
                      const int32_t* __restrict__ ielfc, const double* __restrict__ fcdv,
                      const double* __restrict__ cnwt, const double* __restrict__ rho,
                      double* __restrict__ rwork1, double* __restrict__ fcdm,
                      const DevTimestep* ts) {
  ALE_GATE(ts);
  int e = xcd_index();
  if (e >= nel) return;
  NbFc c = load_nbfc(e, ielel, ielfc);
  rwork1[e] = sumflux_el(e, id1, id2, c.nb, c.fc, fcdv);
  double f[4];
  fluxelvl_el(e, id1, id2, c.nb, c.fc, cnwt, fcdv, rho, f);
  reinterpret_cast<double4*>(fcdm)[e] = make_double4(f[0], f[1], f[2], f[3]);
}

// advectel part 2: ale_kn_sumflux(fcdm)->rwork2, ale_kn_updatebasisel,
// ale_kn_fluxelvl(cnmass,fcdm,e)->flux
__global__ void k_el2(int nel, int id1, int id2, double zerocut, double dencut,
                      const int32_t* __restrict__ ielel, const int32_t* __restrict__ ielfc,
                      const double* __restrict__ fcdm, const double* __restrict__ rwork1,
                      const double* __restrict__ cnmass, const double* __restrict__ en,
                      double* __restrict__ rwork2, double* __restrict__ cutv,
                      double* __restrict__ cutm, double* __restrict__ vpr,
                      double* __restrict__ mpr, double* __restrict__ dpr,
                      double* __restrict__ vol, double* __restrict__ mass,
                      double* __restrict__ rho, double* __restrict__ flux,
                      const DevTimestep* ts) {
  ALE_GATE(ts);
  int e = xcd_index();
  if (e >= nel) return;
  NbFc c = load_nbfc(e, ielel, ielfc);
  double totm = sumflux_el(e, id1, id2, c.nb, c.fc, fcdm);
  rwork2[e] = totm;
  double v0 = vol[e], m0 = mass[e], d0 = rho[e];
  vpr[e] = v0; mpr[e] = m0; dpr[e] = d0;
  cutv[e] = zerocut;
  cutm[e] = v0 * dencut;
  double v1 = v0 + rwork1[e];
  double m1 = m0 + totm;
  vol[e] = v1;
  mass[e] = m1;
  rho[e] = m1 / v1;
  double f[4];
  fluxelvl_el(e, id1, id2, c.nb, c.fc, cnmass, fcdm, en, f);
  reinterpret_cast<double4*>(flux)[e] = make_double4(f[0], f[1], f[2], f[3]);
}

// advectel part 3: ale_kn_updateel (sumflux(flux)->rwork1, e update)
__global__ void k_el3(int nel, int id1, int id2, const int32_t* __restrict__ ielel,
                      const int32_t* __restrict__ ielfc, const double* __restrict__ flux,
                      const double* __restrict__ mpr, const double* __restrict__ mass,
                      const double* __restrict__ cutm, double* __restrict__ rwork1,
                      double* __restrict__ en, const DevTimestep* ts) {
  ALE_GATE(ts);
  int e = xcd_index();
  if (e >= nel) return;
  NbFc c = load_nbfc(e, ielel, ielfc);
  double s = sumflux_el(e, id1, id2, c.nb, c.fc, flux);
  rwork1[e] = s;
  double m1 = mass[e];
  if (m1 > cutm[e]) en[e] = (en[e] * mpr[e] + s) / m1;
}

// advectnd part 1 (element): corner gather of ndu/ndv, cnm0 (rwork3) = cnmass,
// ale_kn_fluxbasisnd, and the corner part of ale_kn_massbasisnd
__global__ void k_nd1(int nel, int id1, int id2, const int32_t* __restrict__ ielnd,
                      const int32_t* __restrict__ ielel, const int32_t* __restrict__ ielfc,
                      const double* __restrict__ ndu, const double* __restrict__ ndv,
                      const double* __restrict__ fcdv, const double* __restrict__ fcdm,
                      double* __restrict__ cnu, double* __restrict__ cnv,
                      double* __restrict__ cnm0, double* __restrict__ cndv,
                      double* __restrict__ cndm, double* __restrict__ cnflux,
                      double* __restrict__ cnmass, const DevTimestep* ts) {
  ALE_GATE(ts);
  int e = xcd_index();
  if (e >= nel) return;
  int4 nd = reinterpret_cast<const int4*>(ielnd)[e];
  reinterpret_cast<double4*>(cnu)[e] =
      make_double4(ndu[nd.x - 1], ndu[nd.y - 1], ndu[nd.z - 1], ndu[nd.w - 1]);
  reinterpret_cast<double4*>(cnv)[e] =
      make_double4(ndv[nd.x - 1], ndv[nd.y - 1], ndv[nd.z - 1], ndv[nd.w - 1]);
  double4 cm = reinterpret_cast<const double4*>(cnmass)[e];
  reinterpret_cast<double4*>(cnm0)[e] = cm;
  NbFc c = load_nbfc(e, ielel, ielfc);
  double fl = 0.0;
  double fl4[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i1 = id1; i1 <= id2; ++i1) {
    int i2 = i1 + 2;
    int e1 = c.nb[i1], e2 = c.nb[i2], s1 = c.fc[i1], s2 = c.fc[i2];
    double w1 = fcdv[4 * e1 + s1], w2 = fcdv[4 * e2 + s2];
    double w3 = fcdm[4 * e1 + s1], w4 = fcdm[4 * e2 + s2];
    if (e1 == e) { w1 = 0.0; w3 = 0.0; }
    if (e2 == e) { w2 = 0.0; w4 = 0.0; }
    w1 = w1 - fcdv[4 * e + i1];
    w2 = w2 - fcdv[4 * e + i2];
    w1 = 0.25 * (w1 - w2);
    cndv[4 * e + i1] = w1;
    cndv[4 * e + i2] = w1;
    w1 = w3 - fcdm[4 * e + i1];
    w2 = w4 - fcdm[4 * e + i2];
    w3 = 0.25 * (w1 - w2);
    cndm[4 * e + i1] = w3;
    cndm[4 * e + i2] = w3;
    w3 = 0.25 * (w1 + w2);
    fl4[0] = fl4[0] + w3; fl4[1] = fl4[1] + w3; fl4[2] = fl4[2] + w3; fl4[3] = fl4[3] + w3;
  }
  (void)fl;
  reinterpret_cast<double4*>(cnflux)[e] = make_double4(fl4[0], fl4[1], fl4[2], fl4[3]);
  reinterpret_cast<double4*>(cnmass)[e] =
      make_double4(cm.x + fl4[0], cm.y + fl4[1], cm.z + fl4[2], cm.w + fl4[3]);
}

// advectnd part 2 (node, ALE-order CSR): ale_kn_initbasisnd,
// ale_kn_calcbasisnd and ale_kn_massbasisnd node sums, ale_kn_cutbasisnd
__global__ void k_nd2(int nnd, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                      double zerocut, double dencut, const double* __restrict__ elv0,
                      const double* __restrict__ elv1, const double* __restrict__ cnm0,
                      const double* __restrict__ cnflux, double* __restrict__ ndv0,
                      double* __restrict__ ndv1, double* __restrict__ ndm0,
                      double* __restrict__ ndm1, double* __restrict__ cutv,
                      double* __restrict__ cutm, const DevTimestep* ts) {
  ALE_GATE(ts);
  int n = xcd_index();
  if (n >= nnd) return;
  double v0 = 0.0, v1 = 0.0, m0 = 0.0;
  int b = ptr[n], end = ptr[n + 1];
  for (int k = b; k < end; ++k) {
    int c = idx[k], e = c >> 2;
    v0 = v0 + 0.25 * elv0[e];
    v1 = v1 + 0.25 * elv1[e];
    m0 = m0 + cnm0[c];
  }
  double m1 = m0;
  for (int k = b; k < end; ++k) m1 = m1 + cnflux[idx[k]];
  ndv0[n] = v0; ndv1[n] = v1; ndm0[n] = m0; ndm1[n] = m1;
  cutv[n] = zerocut;
  cutm[n] = dencut * v0;
}

// ale_kn_fluxndvl for one element (ale_kn_advectors.f90:83-153)
__global__ void k_fluxndvl(int nel, const int32_t* __restrict__ ielel,
                           const int32_t* __restrict__ ielfc, const double* __restrict__ cb,
                           const double* __restrict__ cdb, const double* __restrict__ var,
                           double* __restrict__ cnflux, const DevTimestep* ts) {
  ALE_GATE(ts);
  int e = xcd_index();
  if (e >= nel) return;
  NbFc c = load_nbfc(e, ielel, ielfc);
  double out[4];
  fluxndvl_el(e, c, cb, cdb, var, out);
  reinterpret_cast<double4*>(cnflux)[e] = make_double4(out[0], out[1], out[2], out[3]);
}

// ale_kn_activend + ale_kn_updatend (ale_kn_advectors.f90:182-214,
// ale_kn_advect.f90:240-258): node gather of the corner flux in element order
// base1_out (may be null) receives base1: the post-remap nodal mass is staged
// outside store1 because store1 still holds the element volumes elv0 while
// the node sums run (ale_dr_advect.f90:241-268 reads store1 as elv0, then
// overwrites it with the nodal mass)
__global__ void k_updatend(int nnd, const int32_t* __restrict__ ptr,
                           const int32_t* __restrict__ idx, int ibc,
                           const int32_t* __restrict__ indstatus,
                           const int32_t* __restrict__ indtype, uint8_t* __restrict__ zactive,
                           int write_active, const double* __restrict__ base0,
                           const double* __restrict__ base1, double* __restrict__ base1_out,
                           const double* __restrict__ cut,
                           const double* __restrict__ cnflux, double* __restrict__ ndflux,
                           double* __restrict__ var, const DevTimestep* ts) {
  ALE_GATE(ts);
  int n = xcd_index();
  if (n >= nnd) return;
  double s = 0.0;
  for (int k = ptr[n]; k < ptr[n + 1]; ++k) s = s + cnflux[idx[k]];
  ndflux[n] = s;
  int t = indtype[n];
  bool act = indstatus[n] > 0 && t != ibc && t != -3;
  if (write_active) zactive[n] = act ? 1 : 0;
  double b1 = base1[n];
  if (base1_out) base1_out[n] = b1;
  if (act && b1 > cut[n]) var[n] = (var[n] * base0[n] + s) / b1;
}

// ale_dr_hydroupdate in one pass (ale_dr_update.f90:33-51): geometry on the
// remapped mesh (gather, getiso, getvolume, checkvolume), rho = m/V, EOS
__global__ void k_ale_update(int nel, const int32_t* __restrict__ ielnd,
                             const double* __restrict__ ndx, const double* __restrict__ ndy,
                             double* __restrict__ cnx, double* __restrict__ cny,
                             double* __restrict__ a1, double* __restrict__ a2,
                             double* __restrict__ a3, double* __restrict__ b1,
                             double* __restrict__ b2, double* __restrict__ b3,
                             double* __restrict__ cnwt, double* __restrict__ vol,
                             const double* __restrict__ mass, double* __restrict__ rho,
                             const double* __restrict__ en, double* __restrict__ pre,
                             double* __restrict__ cs2, const int32_t* __restrict__ imat,
                             const int32_t* __restrict__ itype,
                             const double* __restrict__ param, double pcut, double ccut,
                             DtResult* __restrict__ res, const DevTimestep* ts) {
  ALE_GATE(ts);
  int e = xcd_index();
  if (e >= nel) return;
  int4 nd = reinterpret_cast<const int4*>(ielnd)[e];
  double x[4] = {ndx[nd.x - 1], ndx[nd.y - 1], ndx[nd.z - 1], ndx[nd.w - 1]};
  double y[4] = {ndy[nd.x - 1], ndy[nd.y - 1], ndy[nd.z - 1], ndy[nd.w - 1]};
  Iso g;
  getiso(x, y, g);
  reinterpret_cast<double4*>(cnx)[e] = make_double4(x[0], x[1], x[2], x[3]);
  reinterpret_cast<double4*>(cny)[e] = make_double4(y[0], y[1], y[2], y[3]);
  a1[e] = g.a1; a2[e] = g.a2; a3[e] = g.a3;
  b1[e] = g.b1; b2[e] = g.b2; b3[e] = g.b3;
  reinterpret_cast<double4*>(cnwt)[e] = make_double4(g.wt[0], g.wt[1], g.wt[2], g.wt[3]);
  vol[e] = g.vol;
  if (g.vol < 0.0) atomicMax(&res->negvol, e + 1);
  double r = mass[e] / g.vol;
  rho[e] = r;
  int m = imat[e];
  m = m > 0 ? m : 0;
  const double* p = param + BL_NEOSPARAM * m;
  double ee = en[e];
  pre[e] = eos_pressure(itype[m], p, r, ee, pcut);
  cs2[e] = eos_cs2(itype[m], p, r, ee, pcut, ccut);
}

}  // namespace bl

using namespace bl;

extern "C" {

int bl_ale_active(bl_handle_t* h, const bl_timestep_t* ts) {
  bl_touch(h);
  if (!h || !ts || !h->cfg.zexist) return 0;
  return (ts->time >= h->cfg.mintime) && (ts->time <= h->cfg.maxtime);
}

}  // extern "C"

namespace bl {

static int meshstatus(bl_handle_t* h, const DevTimestep* dts) {
  if (h->cfg.zeul || h->relax_type != BL_RELAX_OFF) {
    int nnd = h->sz.nnd;
    BL_LAUNCH(h, "ale_meshstatus", k_meshstatus, nblocks(nnd), kBlock, nnd,
              h->I(BL_ISCRATCH11), dts);
  }
  return BL_SUCCESS;
}

static int fluxvolume(bl_handle_t* h, double dt, const DevTimestep* dts) {
  int nnd = h->sz.nnd, nel = h->sz.nel;
  // the relaxation's mesh velocity replaces the Lagrangian one in store4/5
  if (!h->cfg.zeul && h->relax_type != BL_RELAX_OFF &&
      mesh_motion(h, h->R(BL_NDX), h->R(BL_NDY), h->R(BL_RSCRATCH14), h->R(BL_RSCRATCH15), dt,
                  dts))
    return BL_FAILURE;
  BL_LAUNCH(h, "ale_fluxvol_nodes", k_fluxvol_nodes, nblocks(nnd), kBlock, nnd, h->cfg.zeul,
            dt, h->R(BL_RSCRATCH14), h->R(BL_RSCRATCH15), h->R(BL_RSCRATCH12),
            h->R(BL_RSCRATCH13), h->R(BL_NDX), h->R(BL_NDY), dts);
  BL_LAUNCH(h, "ale_fluxvol_faces", k_fluxvol_faces, nblocks(nel), kBlock, nel,
            h->cfg.zerocut, h->I(BL_IELND), h->R(BL_RSCRATCH12), h->R(BL_RSCRATCH13),
            h->R(BL_NDX), h->R(BL_NDY), h->R(BL_RSCRATCH21), dts);
  return BL_SUCCESS;
}

}  // namespace bl

extern "C" {

int bl_ale_getmeshstatus(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  return meshstatus(h, nullptr);
}

int bl_ale_getfluxvolume(bl_handle_t* h, const bl_timestep_t* ts) {
  bl_touch(h);
  if (!h || !ts) return BL_FAILURE;
  return fluxvolume(h, ts->dt, nullptr);
}

}  // extern "C"

namespace bl {

// lazily allocated nsz-sized temporaries owned by the handle
double* tmp_buffer(Handle* h, int k) {
  if (!h->d_tmp[k]) {
    if (hipMalloc(&h->d_tmp[k], sizeof(double) * std::max<int64_t>(h->sz.nsz, 1)) != hipSuccess) {
      set_error("out of device memory for temporaries");
      return nullptr;
    }
  }
  return h->d_tmp[k];
}

static int advectel(bl_handle_t* h, int id1, int id2, const DevTimestep* dts) {
  int nel = h->sz.nel;
  unsigned nb = nblocks(nel);
  // 0-based face passes
  int a = id1 - 1, b = id2 - 1;
  BL_LAUNCH(h, "ale_el1", k_el1, nb, kBlock, nel, a, b, h->I(BL_IELEL), h->I(BL_IELFC),
            h->R(BL_RSCRATCH21), h->R(BL_CNWT), h->R(BL_ELDENSITY), h->R(BL_RSCRATCH24),
            h->R(BL_RSCRATCH22), dts);
  BL_LAUNCH(h, "ale_el2", k_el2, nb, kBlock, nel, a, b, h->cfg.zerocut, h->cfg.dencut,
            h->I(BL_IELEL), h->I(BL_IELFC), h->R(BL_RSCRATCH22), h->R(BL_RSCRATCH24),
            h->R(BL_CNMASS), h->R(BL_ELENERGY), h->R(BL_RSCRATCH25), h->R(BL_RSCRATCH15),
            h->R(BL_RSCRATCH16), h->R(BL_RSCRATCH11), h->R(BL_RSCRATCH12), h->R(BL_RSCRATCH13),
            h->R(BL_ELVOLUME), h->R(BL_ELMASS), h->R(BL_ELDENSITY), h->R(BL_RSCRATCH23), dts);
  BL_LAUNCH(h, "ale_el3", k_el3, nb, kBlock, nel, a, b, h->I(BL_IELEL), h->I(BL_IELFC),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH12), h->R(BL_ELMASS), h->R(BL_RSCRATCH16),
            h->R(BL_RSCRATCH24), h->R(BL_ELENERGY), dts);
  return BL_SUCCESS;
}

static int advectnd(bl_handle_t* h, int id1, int id2, const DevTimestep* dts) {
  int nel = h->sz.nel, nnd = h->sz.nnd;
  unsigned nbe = nblocks(nel), nbn = nblocks(nnd);
  int a = id1 - 1, b = id2 - 1;
  BL_LAUNCH(h, "ale_nd1", k_nd1, nbe, kBlock, nel, a, b, h->I(BL_IELND), h->I(BL_IELEL),
            h->I(BL_IELFC), h->R(BL_NDU), h->R(BL_NDV), h->R(BL_RSCRATCH21),
            h->R(BL_RSCRATCH22), h->R(BL_RSCRATCH27), h->R(BL_RSCRATCH28), h->R(BL_RSCRATCH26),
            h->R(BL_RSCRATCH24), h->R(BL_RSCRATCH25), h->R(BL_RSCRATCH23), h->R(BL_CNMASS), dts);
  double* ndm1 = tmp_buffer(h, 0);
  if (!ndm1) return BL_FAILURE;
  BL_LAUNCH(h, "ale_nd2", k_nd2, nbn, kBlock, nnd, h->d_ale_ptr, h->d_ale_idx, h->cfg.zerocut,
            h->cfg.dencut, h->R(BL_RSCRATCH11), h->R(BL_ELVOLUME), h->R(BL_RSCRATCH26),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH13), h->R(BL_RSCRATCH14), h->R(BL_RSCRATCH12),
            ndm1, h->R(BL_RSCRATCH15), h->R(BL_RSCRATCH16), dts);
  BL_LAUNCH(h, "ale_fluxndvl", k_fluxndvl, nbe, kBlock, nel, h->I(BL_IELEL), h->I(BL_IELFC),
            h->R(BL_RSCRATCH26), h->R(BL_RSCRATCH25), h->R(BL_RSCRATCH27), h->R(BL_RSCRATCH23),
            dts);
  BL_LAUNCH(h, "ale_updatend", k_updatend, nbn, kBlock, nnd, h->d_el_ptr, h->d_el_idx, -1,
            h->I(BL_ISCRATCH11), h->I(BL_INDTYPE), h->Z(BL_ZSCRATCH11), 0, h->R(BL_RSCRATCH12),
            (const double*)ndm1, h->R(BL_RSCRATCH11), h->R(BL_RSCRATCH16), h->R(BL_RSCRATCH23),
            h->R(BL_RSCRATCH22), h->R(BL_NDU), dts);
  BL_LAUNCH(h, "ale_fluxndvl", k_fluxndvl, nbe, kBlock, nel, h->I(BL_IELEL), h->I(BL_IELFC),
            h->R(BL_RSCRATCH26), h->R(BL_RSCRATCH25), h->R(BL_RSCRATCH28), h->R(BL_RSCRATCH23),
            dts);
  BL_LAUNCH(h, "ale_updatend", k_updatend, nbn, kBlock, nnd, h->d_el_ptr, h->d_el_idx, -2,
            h->I(BL_ISCRATCH11), h->I(BL_INDTYPE), h->Z(BL_ZSCRATCH11), 1, h->R(BL_RSCRATCH12),
            (const double*)h->R(BL_RSCRATCH11), (double*)nullptr, h->R(BL_RSCRATCH16),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH22), h->R(BL_NDV), dts);
  return BL_SUCCESS;
}

// the ALE step inside a captured step: mesh status, flux volumes, unsplit
// advection (adv_type=1) and the hydro update, gated by the device timestep
int ale_step_device(Handle* hb, const DevTimestep* d_ts) {
  bl_handle_t* h = static_cast<bl_handle_t*>(hb);
  if (meshstatus(h, d_ts) || fluxvolume(h, 0.0, d_ts) || advectel(h, 1, 2, d_ts) ||
      advectnd(h, 1, 2, d_ts))
    return BL_FAILURE;
  int nel = h->sz.nel;
  BL_LAUNCH(h, "ale_update", k_ale_update, nblocks(nel), kBlock, nel, h->I(BL_IELND),
            h->R(BL_NDX), h->R(BL_NDY), h->R(BL_CNX), h->R(BL_CNY), h->R(BL_A1), h->R(BL_A2),
            h->R(BL_A3), h->R(BL_B1), h->R(BL_B2), h->R(BL_B3), h->R(BL_CNWT),
            h->R(BL_ELVOLUME), h->R(BL_ELMASS), h->R(BL_ELDENSITY), h->R(BL_ELENERGY),
            h->R(BL_ELPRESSURE), h->R(BL_ELCS2), h->I(BL_IELMAT), h->d_eos_itype,
            h->d_eos_param, h->cfg.pcut, h->cfg.ccut, h->d_dtres, d_ts);
  // mixed cells: component geometry, density and EOS of the hydro update
  if (mix_geometry(h, d_ts) || mix_density(h, d_ts) || mix_eos(h, d_ts)) return BL_FAILURE;
  return BL_SUCCESS;
}

}  // namespace bl

extern "C" {

int bl_ale_advect(bl_handle_t* h, const bl_timestep_t* ts, bl_error_t* err) {
  bl_touch(h);
  if (!h || !ts) return BL_FAILURE;
  if (!h->maps_built) { set_error("bl_ale_advect: maps not built"); return BL_FAILURE; }
  switch (h->cfg.adv_type) {
    case 1:
      if (advectel(h, 1, 2, nullptr) || advectnd(h, 1, 2, nullptr)) return BL_FAILURE;
      break;
    case 2: {
      int ii = (ts->nstep + 1) % 2;
      int i1 = 1 + ii, i2 = 2 - ii, i3 = i2 - i1;
      for (ii = i1; (i3 > 0) ? (ii <= i2) : (ii >= i2); ii += i3)
        if (advectel(h, ii, ii, nullptr) || advectnd(h, ii, ii, nullptr)) return BL_FAILURE;
      break;
    }
    default:
      if (err) {
        err->ierr = BL_FAILURE; err->iout = BL_HALT_ALL;
        std::memset(err->serr, ' ', BL_SLEN);
        const char* m = "ERROR: unrecognised adv_type";
        std::memcpy(err->serr, m, std::strlen(m));
      }
      set_error("unrecognised adv_type");
      return BL_FAILURE;
  }
  if (err) { err->ierr = BL_SUCCESS; err->iout = 0; std::memset(err->serr, ' ', BL_SLEN); }
  return BL_SUCCESS;
}

int bl_ale_hydroupdate(bl_handle_t* h, bl_error_t* err) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int rc = bl_geometry_getgeometry(h, err);
  if (bl_density_getdensity(h) || bl_hydro_geteos(h)) return BL_FAILURE;
  return rc;
}

}  // extern "C"
