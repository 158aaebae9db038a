// bl_api.hip -- C-ABI of the BookLeaf HIP library: device mirror of the dataAPI
// registry, transfers, node->corner maps, initialisation and the Lagrangian
// step drivers (src/packages/hydro/driver/*.f90) as gfx950 kernels.
//
// Layout in HBM: one device array per dataAPI field with the reference's
// shape -- element arrays (nel), corner/face arrays (NCORN,nel) with the 4
// corners of an element contiguous (32 B), node arrays (nnd).  One thread per
// element (or node); a wave64 reads 64 consecutive elements' corner records,
// i.e. 2 KiB contiguous per corner field.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bl_internal.h"
#include "bl_tiles.h"
#include "bl_kernels.h"
#include "bl_mix.h"
#include "bl_math.h"

namespace bl {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// field registry: shapes of src/utilities/data/data.f90:169-430
// ---------------------------------------------------------------------------
static void field_shape(const bl_sizes_t& z, int id, int32_t* kind, int64_t* count) {
  int64_t nel = z.nel2, nnd = z.nnd2, nsz = z.nsz, mcp = z.mcp, mmx = z.mmx;
  *kind = BL_KIND_R64;
  switch (id) {
    case BL_ELDENSITY: case BL_ELENERGY: case BL_ELPRESSURE: case BL_ELCS2:
    case BL_ELVOLUME: case BL_A1: case BL_A2: case BL_A3: case BL_B1:
    case BL_B2: case BL_B3: case BL_ELVISC: case BL_ELMASS:
      *count = nel; break;
    case BL_CNWT: case BL_CNX: case BL_CNY: case BL_CNVISCX: case BL_CNVISCY:
    case BL_CNMASS: case BL_SPMASS:
      *count = 4 * nel; break;
    case BL_RSCRATCH21: case BL_RSCRATCH22: case BL_RSCRATCH23:
    case BL_RSCRATCH24: case BL_RSCRATCH25: case BL_RSCRATCH26:
    case BL_RSCRATCH27: case BL_RSCRATCH28:
      *count = 4 * nsz; break;
    case BL_RSCRATCH11: case BL_RSCRATCH12: case BL_RSCRATCH13:
    case BL_RSCRATCH14: case BL_RSCRATCH15: case BL_RSCRATCH16:
    case BL_RSCRATCH17:
      *count = nsz; break;
    case BL_NDX: case BL_NDY: case BL_NDU: case BL_NDV:
      *count = nnd; break;
    case BL_INDTYPE: case BL_INDLOCGLOB:
      *kind = BL_KIND_I32; *count = nnd; break;
    case BL_IELSORT1: case BL_IELMAT: case BL_IELREG: case BL_IELSORT2:
    case BL_IELLOCGLOB:
      *kind = BL_KIND_I32; *count = nel; break;
    case BL_IELND: case BL_IELEL: case BL_IELFC:
      *kind = BL_KIND_I32; *count = 4 * nel; break;
    case BL_ISCRATCH11:
      *kind = BL_KIND_I32; *count = nsz; break;
    case BL_ZSCRATCH11:
      *kind = BL_KIND_BOOL; *count = nsz; break;
    case BL_CPDENSITY: case BL_CPENERGY: case BL_CPPRESSURE: case BL_CPCS2:
    case BL_CPVOLUME: case BL_FRVOLUME: case BL_CPMASS: case BL_FRMASS:
    case BL_CPVISC: case BL_CPA1: case BL_CPA3: case BL_CPB1: case BL_CPB3:
    case BL_RCPSCRATCH11:
      *count = mcp; break;
    case BL_CPVISCX: case BL_CPVISCY: case BL_RCPSCRATCH21:
    case BL_RCPSCRATCH22: case BL_RCPSCRATCH23: case BL_RCPSCRATCH24:
      *count = 4 * mcp; break;
    case BL_ICPMAT: case BL_ICPNEXT: case BL_ICPPREV: case BL_ICPSCRATCH11:
    case BL_ICPSCRATCH12:
      *kind = BL_KIND_I32; *count = mcp; break;
    case BL_IMXEL: case BL_IMXFCP: case BL_IMXNCP:
      *kind = BL_KIND_I32; *count = mmx; break;
    default:
      *count = 0;
  }
}

static size_t kind_bytes(int32_t kind) {
  return kind == BL_KIND_R64 ? 8 : (kind == BL_KIND_I32 ? 4 : 1);
}

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
static void timing_flush(Handle* h) {
  if (h->ev_used == 0) return;
  hipStreamSynchronize(h->stream);
  for (int s = 0; s < h->ev_used; ++s) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, h->ev_pool[2 * s], h->ev_pool[2 * s + 1]);
    TimingClass& c = h->tclass[h->ev_class[s]];
    c.total_ms += ms;
    c.launches += 1;
  }
  h->ev_used = 0;
}

int timing_begin(Handle* h, const char* cls) {
  int ci = -1;
  for (size_t i = 0; i < h->tclass.size(); ++i)
    if (h->tclass[i].name == cls) { ci = (int)i; break; }
  if (ci < 0) {
    TimingClass c;
    c.name = cls;
    h->tclass.push_back(c);
    ci = (int)h->tclass.size() - 1;
  }
  if (h->ev_used * 2 >= (int)h->ev_pool.size()) timing_flush(h);
  int slot = h->ev_used++;
  h->ev_class[slot] = ci;
  hipEventRecord(h->ev_pool[2 * slot], h->stream);
  return slot;
}

void timing_end(Handle* h, int slot) { hipEventRecord(h->ev_pool[2 * slot + 1], h->stream); }

void invalidate_graph(Handle* h) {
  for (auto& g : h->graph_exec)
    if (g) { hipGraphExecDestroy(g); g = nullptr; }
}

// ---------------------------------------------------------------------------
// element / node kernels of the Lagrangian step
// ---------------------------------------------------------------------------
__global__ void k_fill_r(double* __restrict__ p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i(int32_t* __restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_z(uint8_t* __restrict__ p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_copy_r(const double* __restrict__ a, double* __restrict__ b, int n) {
  int i = xcd_index();
  if (i < n) b[i] = a[i];
}

// geometry_dr_getgeometry: utils_kn_cngather x2, geometry_kn_getiso,
// geometry_kn_getvolume, geometry_kn_checkvolume (geometry_dr_volume.f90:55-95)
__global__ void k_getgeometry(int nel, const int32_t* __restrict__ ielnd,
                              const double* __restrict__ ndx, const double* __restrict__ ndy,
                              double* __restrict__ cnx, double* __restrict__ cny,
                              double* __restrict__ a1, double* __restrict__ a2,
                              double* __restrict__ a3, double* __restrict__ b1,
                              double* __restrict__ b2, double* __restrict__ b3,
                              double* __restrict__ cnwt, double* __restrict__ vol,
                              DtResult* __restrict__ res) {
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
}

// utils_kn_divide (utils_kn_math.f90:28-41)
__global__ void k_divide(int n, const double* __restrict__ num, const double* __restrict__ den,
                         double* __restrict__ out) {
  int i = xcd_index();
  if (i < n) out[i] = num[i] / den[i];
}

// eos_kn_geteos (eos_kn_geteos.f90:28-52)
__global__ void k_geteos(int n, const int32_t* __restrict__ imat, const double* __restrict__ rho,
                         const double* __restrict__ en, double* __restrict__ pre,
                         double* __restrict__ cs2, const int32_t* __restrict__ itype,
                         const double* __restrict__ param, double pcut, double ccut) {
  int i = xcd_index();
  if (i >= n) return;
  int m = imat[i];
  m = m > 0 ? m : 0;
  int t = itype[m];
  const double* p = param + BL_NEOSPARAM * m;
  double r = rho[i], e = en[i];
  pre[i] = eos_pressure(t, p, r, e, pcut);
  cs2[i] = eos_cs2(t, p, r, e, pcut, ccut);
}

// init_kn_elmass (init_kn.f90:32-48)
__global__ void k_elmass(int nel, const double* __restrict__ rho, const double* __restrict__ vol,
                         const double* __restrict__ cnwt, double* __restrict__ m,
                         double* __restrict__ cm) {
  int e = xcd_index();
  if (e >= nel) return;
  double r = rho[e];
  m[e] = r * vol[e];
  double4 w = reinterpret_cast<const double4*>(cnwt)[e];
  reinterpret_cast<double4*>(cm)[e] = make_double4(r * w.x, r * w.y, r * w.z, r * w.w);
}

// hydro_kn_spmassinit (hydro_kn_init.f90:45-77)
__global__ void k_spmassinit(int nel, const double* __restrict__ rho,
                             const double* __restrict__ cnx, const double* __restrict__ cny,
                             double* __restrict__ spm) {
  int e = xcd_index();
  if (e >= nel) return;
  double4 X = reinterpret_cast<const double4*>(cnx)[e];
  double4 Y = reinterpret_cast<const double4*>(cny)[e];
  double x[4] = {X.x, X.y, X.z, X.w}, y[4] = {Y.x, Y.y, Y.z, Y.w}, s[4];
  spmassinit(rho[e], x, y, s);
  reinterpret_cast<double4*>(spm)[e] = make_double4(s[0], s[1], s[2], s[3]);
}

// geometry_kn_getvertex (geometry_kn_position.f90:28-43)
__global__ void k_getvertex(int nnd, double dt, double* __restrict__ ndx,
                            double* __restrict__ ndy, const double* __restrict__ ndu,
                            const double* __restrict__ ndv) {
  int i = xcd_index();
  if (i >= nnd) return;
  ndx[i] = ndx[i] + dt * ndu[i];
  ndy[i] = ndy[i] + dt * ndv[i];
}

// hydro_kn_getenergy (hydro_kn_getenergy.f90:29-52)
__global__ void k_getenergy(int nel, double dt, double zerocut, double* __restrict__ en,
                            const double* __restrict__ m, const double* __restrict__ cnfx,
                            const double* __restrict__ cnfy, const double* __restrict__ cnu,
                            const double* __restrict__ cnv) {
  int e = xcd_index();
  if (e >= nel) return;
  double4 FX = reinterpret_cast<const double4*>(cnfx)[e];
  double4 FY = reinterpret_cast<const double4*>(cnfy)[e];
  double4 U = reinterpret_cast<const double4*>(cnu)[e];
  double4 V = reinterpret_cast<const double4*>(cnv)[e];
  double fx[4] = {FX.x, FX.y, FX.z, FX.w}, fy[4] = {FY.x, FY.y, FY.z, FY.w};
  double u[4] = {U.x, U.y, U.z, U.w}, v[4] = {V.x, V.y, V.z, V.w};
  en[e] = energy_update(en[e], m[e], fx, fy, u, v, dt, zerocut);
}

// hydro_dr_getforce: P + Q [+ SP] [+ HG] (hydro_dr_getforce.f90:40-91)
__global__ void k_getforce(int nel, const double* __restrict__ p, const double* __restrict__ a1,
                           const double* __restrict__ a3, const double* __restrict__ b1,
                           const double* __restrict__ b3, const double* __restrict__ qx,
                           const double* __restrict__ qy, double* __restrict__ cnfx,
                           double* __restrict__ cnfy, int zsp, int zhg, double dt,
                           const int32_t* __restrict__ ielreg, const double* __restrict__ rho,
                           const double* __restrict__ cs2, const double* __restrict__ cnx,
                           const double* __restrict__ cny, const double* __restrict__ spmass,
                           const double* __restrict__ vol, const double* __restrict__ cnu,
                           const double* __restrict__ cnv, const double* __restrict__ kappareg,
                           const double* __restrict__ pmeritreg) {
  int e = xcd_index();
  if (e >= nel) return;
  double fx[4], fy[4];
  forceP(p[e], a1[e], a3[e], b1[e], b3[e], fx, fy);
  double4 QX = reinterpret_cast<const double4*>(qx)[e];
  double4 QY = reinterpret_cast<const double4*>(qy)[e];
  fx[0] = fx[0] + QX.x; fx[1] = fx[1] + QX.y; fx[2] = fx[2] + QX.z; fx[3] = fx[3] + QX.w;
  fy[0] = fy[0] + QY.x; fy[1] = fy[1] + QY.y; fy[2] = fy[2] + QY.z; fy[3] = fy[3] + QY.w;
  if (zsp) {
    int ireg = ielreg[e];
    ireg = ireg < 0 ? -ireg : ireg;
    double4 X = reinterpret_cast<const double4*>(cnx)[e];
    double4 Y = reinterpret_cast<const double4*>(cny)[e];
    double4 S = reinterpret_cast<const double4*>(spmass)[e];
    double x[4] = {X.x, X.y, X.z, X.w}, y[4] = {Y.x, Y.y, Y.z, Y.w};
    double s[4] = {S.x, S.y, S.z, S.w};
    forceSP(pmeritreg[ireg - 1], rho[e], cs2[e], x, y, s, fx, fy);
  }
  if (zhg) {
    double4 U = reinterpret_cast<const double4*>(cnu)[e];
    double4 V = reinterpret_cast<const double4*>(cnv)[e];
    double u[4] = {U.x, U.y, U.z, U.w}, v[4] = {V.x, V.y, V.z, V.w};
    forceHG(kappareg[ielreg[e] - 1], rho[e], vol[e], 1.0 / dt, u, v, fx, fy);
  }
  reinterpret_cast<double4*>(cnfx)[e] = make_double4(fx[0], fx[1], fx[2], fx[3]);
  reinterpret_cast<double4*>(cnfy)[e] = make_double4(fy[0], fy[1], fy[2], fy[3]);
}

// hydro_kn_scatteracceleration as a deterministic node gather through the
// (corner j, element) CSR map -- same per-node summation order as the
// reference's DO jj / DO kk scatter (hydro_kn_getacceleration.f90:30-74)
__global__ void k_scatteracc(int nnd, const int32_t* __restrict__ ptr,
                             const int32_t* __restrict__ idx, double zerocut,
                             const double* __restrict__ rho, const double* __restrict__ cnwt,
                             const double* __restrict__ cnmass, const double* __restrict__ cnfx,
                             const double* __restrict__ cnfy, double* __restrict__ ndarea,
                             double* __restrict__ ndmass, double* __restrict__ ndudot,
                             double* __restrict__ ndvdot) {
  int n = xcd_index();
  if (n >= nnd) return;
  double m = 0.0, a = 0.0, fu = 0.0, fv = 0.0;
  int b = ptr[n], end = ptr[n + 1];
  for (int k = b; k < end; ++k) {
    int c = idx[k];
    int e = c >> 2, j = c & 3;
    double cm = cnmass[c];
    if (cm > zerocut) {
      m = m + cm;
    } else {
      double cp = cnmass[4 * e + ((j + 3) & 3)];
      if (cp > zerocut) m = m + cp;
      else m = m + rho[e] * cnwt[c];
    }
    a = a + cnwt[c];
    fu = fu + cnfx[c];
    fv = fv + cnfy[c];
  }
  ndmass[n] = m;
  ndarea[n] = a;
  ndudot[n] = fu;
  ndvdot[n] = fv;
}

// hydro_kn_getacceleration + utils_kn_bc (hydro_kn_getacceleration.f90:76-100,
// utils_kn_bc.f90:28-58)
__global__ void k_getacc(int nnd, double dencut, double zerocut, double accut2,
                         const double* __restrict__ ndarea, double* __restrict__ ndmass,
                         double* __restrict__ ndu, double* __restrict__ ndv,
                         const int32_t* __restrict__ indtype) {
  int n = xcd_index();
  if (n >= nnd) return;
  double w1 = dencut * ndarea[n];
  double m = ndmass[n], u = ndu[n], v = ndv[n];
  if (m > w1) {
    u = u / m;
    v = v / m;
  } else {
    u = 0.0;
    v = 0.0;
    ndmass[n] = dmax(zerocut, w1);
  }
  int t = indtype[n];
  if (t == -1) u = 0.0;
  else if (t == -2) v = 0.0;
  else if (t == -3) { u = 0.0; v = 0.0; }
  double w2 = u * u + v * v;
  if (w2 < accut2) { u = 0.0; v = 0.0; }
  ndu[n] = u;
  ndv[n] = v;
}

// hydro_kn_applyacceleration (hydro_kn_getacceleration.f90:102-122)
__global__ void k_applyacc(int nnd, double dt, double* __restrict__ ndubar,
                           double* __restrict__ ndvbar, double* __restrict__ ndu,
                           double* __restrict__ ndv) {
  int n = xcd_index();
  if (n >= nnd) return;
  double dt05 = 0.5 * dt;
  double w1 = ndu[n], w2 = ndv[n], au = ndubar[n], av = ndvbar[n];
  ndu[n] = w1 + dt * au;
  ndv[n] = w2 + dt * av;
  ndubar[n] = w1 + dt05 * au;
  ndvbar[n] = w2 + dt05 * av;
}

// utils_kn_cngather of two node arrays (utils_kn_gather.f90:35-51)
__global__ void k_cngather2(int nel, const int32_t* __restrict__ ielnd,
                            const double* __restrict__ a, const double* __restrict__ b,
                            double* __restrict__ ca, double* __restrict__ cb) {
  int e = xcd_index();
  if (e >= nel) return;
  int4 nd = reinterpret_cast<const int4*>(ielnd)[e];
  reinterpret_cast<double4*>(ca)[e] =
      make_double4(a[nd.x - 1], a[nd.y - 1], a[nd.z - 1], a[nd.w - 1]);
  reinterpret_cast<double4*>(cb)[e] =
      make_double4(b[nd.x - 1], b[nd.y - 1], b[nd.z - 1], b[nd.w - 1]);
}

// --- hydro_dr_getdt ---------------------------------------------------------
// (a) corner gather of ndu/ndv, hydro_kn_initeviscosity edge differences and
//     the divergence control hydro_kn_getdtdiv with its max reduction
__global__ void k_getdt_a(int nel, const int32_t* __restrict__ ielnd,
                          const double* __restrict__ ndu, const double* __restrict__ ndv,
                          const double* __restrict__ cnx, const double* __restrict__ cny,
                          const double* __restrict__ a1, const double* __restrict__ a3,
                          const double* __restrict__ b1, const double* __restrict__ b3,
                          const double* __restrict__ vol, double* __restrict__ cnu,
                          double* __restrict__ cnv, double* __restrict__ du,
                          double* __restrict__ dv, double* __restrict__ dx,
                          double* __restrict__ dy, double* __restrict__ part,
                          int32_t* __restrict__ parti) {
  int e = xcd_index();
  double dval = -1.0;  // below TINY: never selected
  if (e < nel) {
    int4 nd = reinterpret_cast<const int4*>(ielnd)[e];
    double u[4] = {ndu[nd.x - 1], ndu[nd.y - 1], ndu[nd.z - 1], ndu[nd.w - 1]};
    double v[4] = {ndv[nd.x - 1], ndv[nd.y - 1], ndv[nd.z - 1], ndv[nd.w - 1]};
    double4 X = reinterpret_cast<const double4*>(cnx)[e];
    double4 Y = reinterpret_cast<const double4*>(cny)[e];
    double x[4] = {X.x, X.y, X.z, X.w}, y[4] = {Y.x, Y.y, Y.z, Y.w};
    reinterpret_cast<double4*>(cnu)[e] = make_double4(u[0], u[1], u[2], u[3]);
    reinterpret_cast<double4*>(cnv)[e] = make_double4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<double4*>(du)[e] =
        make_double4(u[1] - u[0], u[2] - u[1], u[3] - u[2], u[0] - u[3]);
    reinterpret_cast<double4*>(dv)[e] =
        make_double4(v[1] - v[0], v[2] - v[1], v[3] - v[2], v[0] - v[3]);
    reinterpret_cast<double4*>(dx)[e] =
        make_double4(x[1] - x[0], x[2] - x[1], x[3] - x[2], x[0] - x[3]);
    reinterpret_cast<double4*>(dy)[e] =
        make_double4(y[1] - y[0], y[2] - y[1], y[3] - y[2], y[0] - y[3]);
    dval = divterm(u, v, a1[e], a3[e], b1[e], b3[e], vol[e]);
  }
  block_argmax_first(dval, e + 1, part, parti);
}

// (b) hydro_kn_limiteviscosity + hydro_kn_geteviscosity + the CFL length of
//     hydro_kn_getdtcfl (hydro_kn_edgeviscosity.f90:73-282,
//     hydro_kn_getdt.f90:50-66)
__global__ void k_getdt_b(int nel, double zerocut, double cvisc1, double cvisc2, double zcut,
                          const int32_t* __restrict__ indtype, const int32_t* __restrict__ ielel,
                          const int32_t* __restrict__ ielnd, const int32_t* __restrict__ ielfc,
                          const int32_t* __restrict__ ielreg,
                          const int32_t* __restrict__ zdtnotreg,
                          const int32_t* __restrict__ zmidlength, const double* __restrict__ rho,
                          const double* __restrict__ cs2, const double* __restrict__ du,
                          const double* __restrict__ dv, const double* __restrict__ dx,
                          const double* __restrict__ dy, const double* __restrict__ cnx,
                          const double* __restrict__ cny, const double* __restrict__ cnu,
                          const double* __restrict__ cnv, double* __restrict__ rs27,
                          double* __restrict__ cnviscx, double* __restrict__ cnviscy,
                          double* __restrict__ elvisc, double* __restrict__ ellength) {
  int e = xcd_index();
  if (e >= nel) return;
  const int4 nb4 = reinterpret_cast<const int4*>(ielel)[e];
  const int4 fc4 = reinterpret_cast<const int4*>(ielfc)[e];
  const int4 nd4 = reinterpret_cast<const int4*>(ielnd)[e];
  const int nb[4] = {nb4.x - 1, nb4.y - 1, nb4.z - 1, nb4.w - 1};
  const int fc[4] = {fc4.x - 1, fc4.y - 1, fc4.z - 1, fc4.w - 1};
  const int nd[4] = {nd4.x - 1, nd4.y - 1, nd4.z - 1, nd4.w - 1};
  double4 DU = reinterpret_cast<const double4*>(du)[e];
  double4 DV = reinterpret_cast<const double4*>(dv)[e];
  double4 DX = reinterpret_cast<const double4*>(dx)[e];
  double4 DY = reinterpret_cast<const double4*>(dy)[e];
  const double ou[4] = {DU.x, DU.y, DU.z, DU.w}, ov[4] = {DV.x, DV.y, DV.z, DV.w};
  const double ox[4] = {DX.x, DX.y, DX.z, DX.w}, oy[4] = {DY.x, DY.y, DY.z, DY.w};
  const double cs2e = cs2[e], rhoe = rho[e];
  double ev = 0.0;
  double ed[4];
  double sc[4];
#pragma unroll
  for (int iside = 0; iside < 2; ++iside) {  // reference iside = 1,2
    const int is1 = (iside + 3) & 3;          // MOD(iside+2,4)+1, 0-based
    const int is2 = iside + 1;
    const int in1 = nb[iside], in2 = nb[iside + 2];
    // neighbour edges: (MOD(ielfc,4)+1, in1), (MOD(ielfc+2,4)+1, in2) ...
    const int f1a = (fc[iside] + 1) & 3, f2a = (fc[iside + 2] + 3) & 3;
    const int f1b = (fc[iside] + 3) & 3, f2b = (fc[iside + 2] + 1) & 3;
    double s1, s2, s3, s4;
    limiter_ratios(ou[is1], ov[is1], ox[is1], oy[is1], du[4 * in1 + f1a], dv[4 * in1 + f1a],
                   dx[4 * in1 + f1a], dy[4 * in1 + f1a], du[4 * in2 + f2a], dv[4 * in2 + f2a],
                   dx[4 * in2 + f2a], dy[4 * in2 + f2a], zerocut, s1, s2);
    limiter_ratios(ou[is2], ov[is2], ox[is2], oy[is2], du[4 * in1 + f1b], dv[4 * in1 + f1b],
                   dx[4 * in1 + f1b], dy[4 * in1 + f1b], du[4 * in2 + f2b], dv[4 * in2 + f2b],
                   dx[4 * in2 + f2b], dy[4 * in2 + f2b], zerocut, s3, s4);
    // boundary override (hydro_kn_edgeviscosity.f90:148-178)
    if (in1 == e) {
      int ic1 = nd[iside], ic2 = nd[(iside + 1) & 3];
      if (indtype[ic1] < 0 && indtype[ic2] < 0 && in2 != e) { s1 = 1.0; s3 = 1.0; }
      else { s1 = 0.0; s3 = 0.0; }
    }
    if (in2 == e) {
      int ic1 = nd[iside + 2], ic2 = nd[(iside + 3) & 3];
      if (indtype[ic1] < 0 && indtype[ic2] < 0 && in1 != e) { s2 = 1.0; s4 = 1.0; }
      else { s2 = 0.0; s4 = 0.0; }
    }
    int ins = nb[is1];
    ed[is1] = limiter_edge(s1, s2, ou[is1], ov[is1], cs2e, cs2[ins], rhoe, rho[ins], cvisc1,
                           cvisc2, zerocut, ev);
    ins = nb[is2];
    ed[is2] = limiter_edge(s3, s4, ou[is2], ov[is2], cs2e, cs2[ins], rhoe, rho[ins], cvisc1,
                           cvisc2, zerocut, ev);
    sc[0] = s1; sc[1] = s2; sc[2] = s3; sc[3] = s4;
  }
  reinterpret_cast<double4*>(rs27)[e] = make_double4(sc[0], sc[1], sc[2], sc[3]);
  elvisc[e] = ev;
  double4 X = reinterpret_cast<const double4*>(cnx)[e];
  double4 Y = reinterpret_cast<const double4*>(cny)[e];
  double4 U = reinterpret_cast<const double4*>(cnu)[e];
  double4 V = reinterpret_cast<const double4*>(cnv)[e];
  double x[4] = {X.x, X.y, X.z, X.w}, y[4] = {Y.x, Y.y, Y.z, Y.w};
  double u[4] = {U.x, U.y, U.z, U.w}, v[4] = {V.x, V.y, V.z, V.w};
  double qx[4], qy[4];
  geteviscosity(x, y, u, v, ed, ed, zerocut, qx, qy);
  reinterpret_cast<double4*>(cnviscx)[e] = make_double4(qx[0], qx[1], qx[2], qx[3]);
  reinterpret_cast<double4*>(cnviscy)[e] = make_double4(qy[0], qy[1], qy[2], qy[3]);
  int ireg = ielreg[e] - 1;
  ellength[e] = zdtnotreg[ireg] ? DBL_MIN : minlength(x, y, zcut, zmidlength[ireg] != 0);
}

// (c) hydro_dr_getcs2 (cs2 += elvisc) and the CFL ratio with its min reduction
//     (hydro_kn_getdt.f90:50-73)
__global__ void k_getdt_c(int nel, const int32_t* __restrict__ ielreg,
                          const int32_t* __restrict__ zdtnotreg,
                          const double* __restrict__ elvisc, double* __restrict__ cs2,
                          const double* __restrict__ ellength, double* __restrict__ rs11,
                          double* __restrict__ part, int32_t* __restrict__ parti, int nblk_a,
                          const int32_t* __restrict__ ielmat, MixArgs M) {
  int e = xcd_index();
  double r = DBL_MAX;
  if (e < nel) {
    double c = cs2[e] + elvisc[e];
    // mixed cell: the frmass average of the components (hydro_dr_getcs2.f90:40-42)
    if (M.ncp > 0 && ielmat[e] < 0) c = mix_cs2(M, -ielmat[e] - 1);
    cs2[e] = c;
    r = zdtnotreg[ielreg[e] - 1] ? DBL_MAX : ellength[e] / c;
    rs11[e] = r;
  }
  block_argmin_first(r, e + 1, part + 2 * nblk_a, parti + 2 * nblk_a);
}

// (d) one workgroup combines the per-workgroup partials of (a) and (c); a
//     separate launch instead of a "last workgroup" ticket, which would need
//     a device-scope fence in every workgroup (an L2 write-back per XCD)
__global__ void k_dt_final(int nblk_a, int nblk_c, const double* __restrict__ part,
                           const int32_t* __restrict__ parti, DtResult* __restrict__ res) {
  final_dt_reduce(nblk_a, nblk_c, part, parti, res);
}

// --- the allocation-free timestep (bl_hydro_getdt_noalloc) ------------------
// Its workspace, the partials of k_getdt_a/k_getdt_c and k_dt_final's
// result, lives in this file-scope array: one copy per GPU, held until the
// process ends.  Layout: the DtResult in the first 256 bytes, then the
// partial values part[3 * kNaBlocks] (DIV in [0, nb), CFL in [2nb, 3nb), as
// the handle's partials), then their indices parti[3 * kNaBlocks].  Every
// workgroup writes its own slots and k_dt_final writes every field the host
// reads, so nothing in it needs clearing between calls.
constexpr int kNaBlocks = BL_NOALLOC_MAX_EL / kBlock;
constexpr size_t kNaPart = 256;
constexpr size_t kNaParti = kNaPart + sizeof(double) * 3 * (size_t)kNaBlocks;
constexpr size_t kNaBytes = kNaParti + sizeof(int32_t) * 3 * (size_t)kNaBlocks;
static_assert(sizeof(DtResult) <= kNaPart, "DtResult does not fit its slot");
static_assert(BL_NOALLOC_MAX_EL % kBlock == 0, "whole workgroups");
static __device__ __align__(256) unsigned char scratch[kNaBytes];

// scratch's device address on each GPU (hipGetSymbolAddress with that GPU
// current; the host-side name is no device address)
constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mutex;
static unsigned char* g_na_ptr[kNaMaxDevices] = {};
static bool g_na_looked[kNaMaxDevices] = {};

// Looks scratch up on the current device (once per device); called by
// bl_create, so the entry point itself makes no runtime look-up.  Returns
// the address, or nullptr with the error set.
static unsigned char* noalloc_scratch(int device) {
  if (device < 0 || device >= kNaMaxDevices) {
    set_error("bl_hydro_getdt_noalloc: device ordinal beyond the workspace table");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_looked[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      set_error(std::string("bl_hydro_getdt_noalloc: hipGetSymbolAddress: ") +
                hipGetErrorString(e));
      return nullptr;
    }
    g_na_ptr[device] = static_cast<unsigned char*>(ptr);
    g_na_looked[device] = true;
  }
  return g_na_ptr[device];
}

}  // namespace bl

using namespace bl;

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* bl_last_error(void) { return g_last_error.c_str(); }

static int fill_field(bl_handle_t* h, int id) {
  int64_t n = h->count[id];
  if (n <= 0) return BL_SUCCESS;
  if (h->kind[id] == BL_KIND_R64)
    BL_LAUNCH(h, "fill", k_fill_r, nblocks(n), kBlock, h->R(id), n, -2.0e12);
  else if (h->kind[id] == BL_KIND_I32)
    BL_LAUNCH(h, "fill", k_fill_i, nblocks(n), kBlock, h->I(id), n, (int32_t)-2000000000);
  else
    BL_LAUNCH(h, "fill", k_fill_z, nblocks(n), kBlock, h->Z(id), n, (uint8_t)0);
  return BL_SUCCESS;
}

int bl_create(const bl_sizes_t* sizes, bl_handle_t** out) {
  if (!sizes || !out) { set_error("bl_create: null argument"); return BL_FAILURE; }
  if (sizes->nel <= 0 || sizes->nnd <= 0) { set_error("bl_create: empty mesh"); return BL_FAILURE; }
  bl_handle_t* h = new bl_handle_t();
  h->sz = *sizes;
  if (h->sz.nel1 <= 0) h->sz.nel1 = h->sz.nel;
  if (h->sz.nel2 <= 0) h->sz.nel2 = h->sz.nel;
  if (h->sz.nnd1 <= 0) h->sz.nnd1 = h->sz.nnd;
  if (h->sz.nnd2 <= 0) h->sz.nnd2 = h->sz.nnd;
  h->sz.nsz = std::max(h->sz.nel2, h->sz.nnd2);  // data.f90:165
  if (hipGetDevice(&h->device) != hipSuccess) { set_error("bl_create: no HIP device"); delete h; return BL_FAILURE; }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("bl_create: stream creation failed"); delete h; return BL_FAILURE;
  }
  for (int id = 1; id <= BL_NFIELD; ++id) {
    field_shape(h->sz, id, &h->kind[id], &h->count[id]);
    int64_t n = std::max<int64_t>(h->count[id], 1);
    if (hipMalloc(&h->f[id], n * kind_bytes(h->kind[id])) != hipSuccess) {
      set_error("bl_create: out of device memory"); bl_destroy(h); return BL_FAILURE;
    }
    if (fill_field(h, id) != BL_SUCCESS) { bl_destroy(h); return BL_FAILURE; }
  }
  h->npart = nblocks(h->sz.nsz) + 16;
  if (hipMalloc(&h->d_part, sizeof(double) * 4 * h->npart) != hipSuccess ||
      hipMalloc(&h->d_parti, sizeof(int32_t) * 4 * h->npart) != hipSuccess ||
      hipMalloc(&h->d_ticket, sizeof(uint32_t) * 16) != hipSuccess ||
      hipMalloc(&h->d_dtres, sizeof(DtResult)) != hipSuccess ||
      hipHostMalloc(&h->h_dtres, sizeof(DtResult)) != hipSuccess ||
      hipMalloc(&h->d_ts, sizeof(DevTimestep)) != hipSuccess ||
      hipHostMalloc(&h->h_ts, sizeof(DevTimestep)) != hipSuccess ||
      hipHostMalloc(&h->h_pub_ts, sizeof(DevTimestep), hipHostMallocCoherent) != hipSuccess) {
    set_error("bl_create: out of device memory"); bl_destroy(h); return BL_FAILURE;
  }
  // the allocation-free timestep's workspace address on this device (a
  // failed look-up is reported by that entry point, not here)
  if (!noalloc_scratch(h->device)) g_last_error.clear();
  hipMemsetAsync(h->d_ticket, 0, sizeof(uint32_t) * 16, h->stream);
  hipMemsetAsync(h->d_dtres, 0, sizeof(DtResult), h->stream);
  hipMemsetAsync(h->d_ts, 0, sizeof(DevTimestep), h->stream);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    set_error("bl_create: initialisation failed"); bl_destroy(h); return BL_FAILURE;
  }
  // defaults of *_cf_defaults (global_config.f90:37-40, hydro_config.f90:43-49,
  // eos_config.f90:42-43, time_config.f90:37-42, ale_config.f90:42-49)
  bl_config_t c{};
  c.zcut = 1.0e-8; c.zerocut = 1.0e-40; c.dencut = 1.0e-6; c.accut = 1.0e-6;
  c.cvisc1 = 0.5; c.cvisc2 = 0.75; c.cfl_sf = 0.5; c.div_sf = 0.25;
  c.ccut = 1.0e-6; c.pcut = 1.0e-8;
  c.time_start = 0.0; c.time_end = 1.0; c.dt_g = 1.02; c.dt_min = 1.0e-8;
  c.dt_max = 1.0e-1; c.dt_initial = 1.0e-5;
  c.mintime = DBL_MAX; c.maxtime = DBL_MAX; c.sf = 0.5; c.adv_type = 1;
  h->cfg = c;
  *out = h;
  return BL_SUCCESS;
}

int bl_destroy(bl_handle_t* h) {
  if (!h) return BL_SUCCESS;
  if (h->stream) hipStreamSynchronize(h->stream);
  invalidate_graph(h);
  const bool sets = h->nodeset[1][0] != nullptr;
  const bool elsets = h->elset[0][0] != nullptr;
  for (int id = 1; id <= BL_NFIELD; ++id) {
    bool node = id == BL_NDX || id == BL_NDY || id == BL_NDU || id == BL_NDV;
    bool state = id == BL_ELDENSITY || id == BL_ELENERGY || id == BL_ELPRESSURE ||
                 id == BL_ELCS2;
    if (h->f[id] && !(sets && node) && !(elsets && state)) hipFree(h->f[id]);
  }
  if (sets)
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < 4; ++k) hipFree(h->nodeset[s][k]);
  if (elsets)
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < 4; ++k) hipFree(h->elset[s][k]);
  free_tiles(h);
  void* bufs[] = {h->d_kappareg, h->d_pmeritreg, h->d_zdtnotreg, h->d_zmidlength,
                  h->d_eos_itype, h->d_eos_param, h->d_hyd_ptr, h->d_hyd_idx,
                  h->d_ale_ptr, h->d_ale_idx, h->d_el_ptr, h->d_el_idx, h->d_part,
                  h->d_parti, h->d_ticket, h->d_dtres, h->d_ts, h->d_hist, h->d_outer,
                  h->d_diag_hist, h->d_diag_count, h->d_diag_part, h->d_ring, h->d_work, h->d_rest, h->d_vstats,
                  h->d_rlist, h->d_zlist, h->d_lflag, h->d_lmark, h->d_ld1, h->d_nbfc, h->d_cw[0], h->d_cw[1],
                  h->d_eoff, h->d_lboff, h->d_einst, h->d_ninst, h->d_inst, h->d_inst_dt, h->d_cnt,
                  h->d_epart, h->d_eparti, h->d_ehist};
  for (void* p : bufs) if (p) hipFree(p);
  for (double* p : h->d_tmp) if (p) hipFree(p);
  for (double* p : h->cterm) if (p) hipFree(p);
  for (double* p : h->ale_tmp) if (p) hipFree(p);
  for (double* p : h->relax_tmp) if (p) hipFree(p);
  if (h->h_dtres) hipHostFree(h->h_dtres);
  if (h->h_ts) hipHostFree(h->h_ts);
  if (h->h_pub_ts) hipHostFree(h->h_pub_ts);
  if (h->h_pub_hist) hipHostFree(h->h_pub_hist);
  for (hipEvent_t ev : h->ev_pool) hipEventDestroy(ev);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return BL_SUCCESS;
}


int bl_set_config(bl_handle_t* h, const bl_config_t* cfg) {
  bl_touch(h);
  if (!h || !cfg) { set_error("bl_set_config: null argument"); return BL_FAILURE; }
  h->cfg = *cfg;
  invalidate_graph(h);
  return BL_SUCCESS;
}

}  // extern "C"

template <typename T>
static int upload_table(bl_handle_t* h, T** dptr, const std::vector<T>& v) {
  if (*dptr) { hipFree(*dptr); *dptr = nullptr; }
  size_t n = std::max<size_t>(v.size(), 1);
  BL_HIP(hipMalloc(dptr, n * sizeof(T)));
  if (!v.empty())
    BL_HIP(hipMemcpyAsync(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  return BL_SUCCESS;
}

extern "C" {

int bl_set_regions(bl_handle_t* h, int32_t nreg, const double* kappareg,
                   const double* pmeritreg, const int32_t* zdtnotreg,
                   const int32_t* zmidlength) {
  bl_touch(h);
  if (!h || nreg <= 0) { set_error("bl_set_regions: bad argument"); return BL_FAILURE; }
  h->nreg = nreg;
  h->kappareg.assign(kappareg, kappareg + nreg);
  h->pmeritreg.assign(pmeritreg, pmeritreg + nreg);
  h->zdtnotreg.assign(zdtnotreg, zdtnotreg + nreg);
  h->zmidlength.assign(zmidlength, zmidlength + nreg);
  if (upload_table(h, &h->d_kappareg, h->kappareg) || upload_table(h, &h->d_pmeritreg, h->pmeritreg) ||
      upload_table(h, &h->d_zdtnotreg, h->zdtnotreg) || upload_table(h, &h->d_zmidlength, h->zmidlength))
    return BL_FAILURE;
  invalidate_graph(h);
  return BL_SUCCESS;
}

int bl_set_eos(bl_handle_t* h, int32_t nmat, const int32_t* itype, const double* param) {
  bl_touch(h);
  if (!h || nmat <= 0) { set_error("bl_set_eos: bad argument"); return BL_FAILURE; }
  h->nmat = nmat;
  h->eos_itype.assign(itype, itype + nmat + 1);
  h->eos_param.assign(param, param + BL_NEOSPARAM * (nmat + 1));
  h->ideal_single = true;
  for (int m = 1; m <= nmat; ++m) if (itype[m] != 1) h->ideal_single = false;
  if (upload_table(h, &h->d_eos_itype, h->eos_itype) || upload_table(h, &h->d_eos_param, h->eos_param))
    return BL_FAILURE;
  invalidate_graph(h);
  return BL_SUCCESS;
}

int bl_field_info(bl_handle_t* h, int32_t id, int32_t* kind, int64_t* count) {
  if (!h || id < 1 || id > BL_NFIELD) { set_error("bl_field_info: bad id"); return BL_FAILURE; }
  if (kind) *kind = h->kind[id];
  if (count) *count = h->count[id];
  return BL_SUCCESS;
}

void* bl_field_device_ptr(bl_handle_t* h, int32_t id) {
  if (!h || id < 1 || id > BL_NFIELD) return nullptr;
  // stores through it are not seen here: no run call resumes after this
  h->ptr_out = true;
  h->resumable = false;
  return h->f[id];
}

int bl_upload(bl_handle_t* h, int32_t id, const void* host, int64_t count) {
  bl_touch(h);
  if (!h || id < 1 || id > BL_NFIELD || !host) { set_error("bl_upload: bad argument"); return BL_FAILURE; }
  if (count < 0 || count > h->count[id]) { set_error("bl_upload: count exceeds field size"); return BL_FAILURE; }
  if (count == 0) return BL_SUCCESS;
  BL_HIP(hipMemcpyAsync(h->f[id], host, count * kind_bytes(h->kind[id]), hipMemcpyHostToDevice, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  if (id == BL_IELND || id == BL_IELSORT1 || id == BL_IELSORT2) { h->maps_built = false; invalidate_graph(h); }
  if (id == BL_IELEL || id == BL_IELFC) { h->nbfc_stale = true; invalidate_graph(h); }
  return BL_SUCCESS;
}

int bl_download(bl_handle_t* h, int32_t id, void* host, int64_t count) {
  if (!h || id < 1 || id > BL_NFIELD || !host) { set_error("bl_download: bad argument"); return BL_FAILURE; }
  if (count < 0 || count > h->count[id]) { set_error("bl_download: count exceeds field size"); return BL_FAILURE; }
  if (count == 0) return BL_SUCCESS;
  BL_HIP(hipMemcpyAsync(host, h->f[id], count * kind_bytes(h->kind[id]), hipMemcpyDeviceToHost, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  return BL_SUCCESS;
}

int bl_synchronize(bl_handle_t* h) {
  if (!h) return BL_FAILURE;
  BL_HIP(hipStreamSynchronize(h->stream));
  return BL_SUCCESS;
}

// counting sort of the (node, corner) incidences in iteration order
static int build_csr(bl_handle_t* h, const std::vector<int32_t>& nodes,
                     const std::vector<int32_t>& corners, int32_t** dptr, int32_t** didx) {
  int nnd = h->sz.nnd;
  std::vector<int32_t> ptr(nnd + 1, 0), idx(nodes.size());
  for (int32_t n : nodes) {
    if (n < 1 || n > nnd) { set_error("bl_build_maps: ielnd entry out of range"); return BL_FAILURE; }
    ptr[n] += 1;
  }
  for (int i = 0; i < nnd; ++i) ptr[i + 1] += ptr[i];
  std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
  for (size_t k = 0; k < nodes.size(); ++k) idx[pos[nodes[k] - 1]++] = corners[k];
  if (*dptr) hipFree(*dptr);
  if (*didx) hipFree(*didx);
  *dptr = nullptr; *didx = nullptr;
  if (upload_table(h, dptr, ptr) || upload_table(h, didx, idx)) return BL_FAILURE;
  return BL_SUCCESS;
}

// The limiter of hydro_kn_getviscosity (hydro_kn_edgeviscosity.f90:120-186)
// reads, for direction iside (0,1) of element e with neighbours in1 =
// ielel(iside) and in2 = ielel(iside+2) and faces f1 = ielfc(iside),
// f2 = ielfc(iside+2), four edges of the neighbours:
//   in1 edge f1+1 = (in1 node f1+1, in1 node f1+2)
//   in1 edge f1+3 = (in1 node f1+3, in1 node f1)
//   in2 edge f2+3 = (in2 node f2+3, in2 node f2)
//   in2 edge f2+1 = (in2 node f2+1, in2 node f2+2)
// When the neighbour traverses the shared face in the opposite sense, one end
// of each is a node of e (nd(iside), nd(iside+1), nd(iside+3), nd(iside+2)
// in that order), and the other, stored here as outer(e, 4*iside + k), is a
// node of the neighbour only.  A boundary side (in == e) has its ratios
// replaced, so any valid node will do there.  Elements that break the
// pattern leave d_outer unset and the fused step uses the general gather.
static int build_outer(bl_handle_t* h, const std::vector<int32_t>& ielnd) {
  if (h->d_outer) hipFree(h->d_outer);
  h->d_outer = nullptr;
  const int nel = h->sz.nel;
  std::vector<int32_t> ielel(4 * (size_t)nel), ielfc(4 * (size_t)nel);
  if (bl_download(h, BL_IELEL, ielel.data(), 4 * (int64_t)nel) ||
      bl_download(h, BL_IELFC, ielfc.data(), 4 * (int64_t)nel))
    return BL_FAILURE;
  std::vector<int32_t> outer(8 * (size_t)nel);
  auto node = [&](int el, int c) { return ielnd[4 * (size_t)el + (c & 3)]; };
  for (int e = 0; e < nel; ++e) {
    for (int is = 0; is < 2; ++is) {
      int32_t* o = &outer[8 * (size_t)e + 4 * is];
      const int in1 = ielel[4 * (size_t)e + is] - 1, in2 = ielel[4 * (size_t)e + is + 2] - 1;
      const int f1 = ielfc[4 * (size_t)e + is] - 1, f2 = ielfc[4 * (size_t)e + is + 2] - 1;
      if (in1 < 0 || in1 >= nel || in2 < 0 || in2 >= nel) return BL_SUCCESS;
      if (in1 == e) {
        o[0] = node(e, is); o[1] = node(e, is + 1);
      } else {
        if (f1 < 0 || f1 > 3 || node(in1, f1 + 1) != node(e, is) ||
            node(in1, f1) != node(e, is + 1))
          return BL_SUCCESS;
        o[0] = node(in1, f1 + 2); o[1] = node(in1, f1 + 3);
      }
      if (in2 == e) {
        o[2] = node(e, is + 3); o[3] = node(e, is + 2);
      } else {
        if (f2 < 0 || f2 > 3 || node(in2, f2) != node(e, is + 3) ||
            node(in2, f2 + 1) != node(e, is + 2))
          return BL_SUCCESS;
        o[2] = node(in2, f2 + 3); o[3] = node(in2, f2 + 2);
      }
    }
  }
  return upload_table(h, &h->d_outer, outer);
}

int bl_build_maps(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nel = h->sz.nel;
  std::vector<int32_t> ielnd(4 * (size_t)nel), s1(nel), s2(nel);
  if (bl_download(h, BL_IELND, ielnd.data(), 4 * (int64_t)nel) ||
      bl_download(h, BL_IELSORT1, s1.data(), nel) || bl_download(h, BL_IELSORT2, s2.data(), nel))
    return BL_FAILURE;
  for (int k = 0; k < nel; ++k) {
    if (s1[k] < 1 || s1[k] > nel) s1[k] = k + 1;  // unset: identity (init_dr_utils.f90:64-69)
    if (s2[k] < 1 || s2[k] > nel) s2[k] = k + 1;
  }
  std::vector<int32_t> nodes, corners;
  nodes.reserve(4 * (size_t)nel);
  corners.reserve(4 * (size_t)nel);
  // hydro order: DO jj=1,4; DO kk=1,nel; iel=ielsort1(kk)
  for (int jj = 0; jj < 4; ++jj)
    for (int kk = 0; kk < nel; ++kk) {
      int iel = s1[kk] - 1;
      nodes.push_back(ielnd[4 * (size_t)iel + jj]);
      corners.push_back(4 * iel + jj);
    }
  if (build_csr(h, nodes, corners, &h->d_hyd_ptr, &h->d_hyd_idx)) return BL_FAILURE;
  // ALE order: DO ii=1,nel; iel=ielsort2(ii); corners 1..4
  nodes.clear(); corners.clear();
  for (int ii = 0; ii < nel; ++ii) {
    int iel = s2[ii] - 1;
    for (int j = 0; j < 4; ++j) { nodes.push_back(ielnd[4 * (size_t)iel + j]); corners.push_back(4 * iel + j); }
  }
  if (build_csr(h, nodes, corners, &h->d_ale_ptr, &h->d_ale_idx)) return BL_FAILURE;
  // plain element order (ale_kn_updatend.f90 loop)
  nodes.clear(); corners.clear();
  for (int iel = 0; iel < nel; ++iel)
    for (int j = 0; j < 4; ++j) { nodes.push_back(ielnd[4 * (size_t)iel + j]); corners.push_back(4 * iel + j); }
  if (build_csr(h, nodes, corners, &h->d_el_ptr, &h->d_el_idx)) return BL_FAILURE;
  if (build_outer(h, ielnd)) return BL_FAILURE;
  if (h->relax_type != BL_RELAX_OFF && relax_rebuild(h)) return BL_FAILURE;
  h->nbfc_stale = true;
  h->tile_state = 0;  // the tiling follows the connectivity
  h->maps_built = true;
  invalidate_graph(h);
  return BL_SUCCESS;
}

int bl_init_mesh(bl_handle_t* h) {
  bl_touch(h);
  // init_kn_connectivity (init_kn.f90:82-91 via utils_kn_connectivity.f90),
  // init_kn_nodetype (init_kn.f90:93-113), ielsort = identity
  // (init_dr_utils.f90:64-69)
  if (!h) return BL_FAILURE;
  int nel = h->sz.nel, nnd = h->sz.nnd;
  std::vector<int32_t> ielnd(4 * (size_t)nel), indtype(nnd);
  if (bl_download(h, BL_IELND, ielnd.data(), 4 * (int64_t)nel) ||
      bl_download(h, BL_INDTYPE, indtype.data(), nnd))
    return BL_FAILURE;
  std::vector<int32_t> ielel(4 * (size_t)nel, 0), ielfc(4 * (size_t)nel, 0);
  // face f of element e joins corners f and f+1; key the unordered node pair
  std::vector<std::pair<uint64_t, int32_t>> keys;
  keys.reserve(4 * (size_t)nel);
  for (int e = 0; e < nel; ++e)
    for (int f = 0; f < 4; ++f) {
      uint64_t a = (uint32_t)ielnd[4 * (size_t)e + f], b = (uint32_t)ielnd[4 * (size_t)e + ((f + 1) & 3)];
      uint64_t k = (std::min(a, b) << 32) | std::max(a, b);
      keys.push_back({k, 4 * e + f});
    }
  std::sort(keys.begin(), keys.end());
  for (size_t i = 0; i + 1 < keys.size(); ++i) {
    if (keys[i].first != keys[i + 1].first) continue;
    int c1 = keys[i].second, c2 = keys[i + 1].second;
    ielel[c1] = c2 / 4 + 1;
    ielel[c2] = c1 / 4 + 1;
  }
  // side connectivity: face of the neighbour that points back (last match,
  // utils_kn_getsconn), then boundary correction (utils_kn_corrconn)
  for (int e = 0; e < nel; ++e)
    for (int f = 0; f < 4; ++f) {
      int nb = ielel[4 * (size_t)e + f];
      if (nb > 0) {
        int s = 0;
        for (int g = 0; g < 4; ++g) if (ielel[4 * (size_t)(nb - 1) + g] == e + 1) s = g + 1;
        ielfc[4 * (size_t)e + f] = s;
      }
    }
  for (int e = 0; e < nel; ++e)
    for (int f = 0; f < 2; ++f) {
      int f2 = f + 2;
      if (ielel[4 * (size_t)e + f] == 0) { ielel[4 * (size_t)e + f] = e + 1; ielfc[4 * (size_t)e + f] = f2 + 1; }
      if (ielel[4 * (size_t)e + f2] == 0) { ielel[4 * (size_t)e + f2] = e + 1; ielfc[4 * (size_t)e + f2] = f + 1; }
    }
  // node types: a corner node of an element with 3 boundary nodes whose
  // neighbours are -1/-2 becomes -3
  for (int e = 0; e < nel; ++e) {
    int nodes[4];
    int cnt = 0;
    for (int j = 0; j < 4; ++j) { nodes[j] = ielnd[4 * (size_t)e + j]; if (indtype[nodes[j] - 1] < 0) ++cnt; }
    if (cnt != 3) continue;
    int ii = 0;
    for (ii = 0; ii < 4; ++ii) if (indtype[nodes[ii] - 1] > 0) break;
    ii = (ii + 2) % 4;
    int jj = nodes[ii];
    if (jj <= nnd) {
      int j1 = nodes[(ii + 1) % 4], j2 = nodes[(ii + 3) % 4];
      if ((indtype[j1 - 1] == -2 && indtype[j2 - 1] == -1) || (indtype[j2 - 1] == -2 && indtype[j1 - 1] == -1))
        indtype[jj - 1] = -3;
    }
  }
  std::vector<int32_t> ident(nel);
  for (int e = 0; e < nel; ++e) ident[e] = e + 1;
  if (bl_upload(h, BL_IELEL, ielel.data(), 4 * (int64_t)nel) ||
      bl_upload(h, BL_IELFC, ielfc.data(), 4 * (int64_t)nel) ||
      bl_upload(h, BL_INDTYPE, indtype.data(), nnd) ||
      bl_upload(h, BL_IELSORT1, ident.data(), nel) || bl_upload(h, BL_IELSORT2, ident.data(), nel))
    return BL_FAILURE;
  return bl_build_maps(h);
}

// --------------------------------------------------------------- utilities
int bl_geometry_getgeometry(bl_handle_t* h, bl_error_t* err) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nel = h->sz.nel;
  BL_HIP(hipMemsetAsync(&h->d_dtres->negvol, 0, sizeof(int32_t), h->stream));
  BL_LAUNCH(h, "getgeometry", k_getgeometry, nblocks(nel), kBlock, nel, h->I(BL_IELND),
            h->R(BL_NDX), h->R(BL_NDY), h->R(BL_CNX), h->R(BL_CNY), h->R(BL_A1), h->R(BL_A2),
            h->R(BL_A3), h->R(BL_B1), h->R(BL_B2), h->R(BL_B3), h->R(BL_CNWT),
            h->R(BL_ELVOLUME), h->d_dtres);
  if (mix_geometry(h, nullptr)) return BL_FAILURE;
  BL_HIP(hipMemcpyAsync(&h->h_dtres->negvol, &h->d_dtres->negvol, sizeof(int32_t),
                        hipMemcpyDeviceToHost, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  int bad = h->h_dtres->negvol;
  if (bad > 0) {
    if (err) {
      err->ierr = -bad;
      err->iout = BL_HALT_SINGLE;
      std::memset(err->serr, ' ', BL_SLEN);
      const char* m = "ERROR: negative volume in geometry_dr_getgeometry";
      std::memcpy(err->serr, m, std::strlen(m));
    }
    set_error("negative volume in geometry_dr_getgeometry");
    return BL_FAILURE;
  }
  if (err) { err->ierr = BL_SUCCESS; err->iout = 0; std::memset(err->serr, ' ', BL_SLEN); }
  return BL_SUCCESS;
}

int bl_density_getdensity(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nel = h->sz.nel;
  BL_LAUNCH(h, "getdensity", k_divide, nblocks(nel), kBlock, nel, h->R(BL_ELMASS),
            h->R(BL_ELVOLUME), h->R(BL_ELDENSITY));
  return mix_density(h, nullptr);
}

int bl_hydro_geteos(bl_handle_t* h) {
  bl_touch(h);
  if (!h || !h->d_eos_itype) { set_error("bl_hydro_geteos: EOS not set"); return BL_FAILURE; }
  int nel = h->sz.nel;
  BL_LAUNCH(h, "geteos", k_geteos, nblocks(nel), kBlock, nel, h->I(BL_IELMAT),
            h->R(BL_ELDENSITY), h->R(BL_ELENERGY), h->R(BL_ELPRESSURE), h->R(BL_ELCS2),
            h->d_eos_itype, h->d_eos_param, h->cfg.pcut, h->cfg.ccut);
  return mix_eos(h, nullptr);
}

int bl_init_state(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nel = h->sz.nel;
  BL_LAUNCH(h, "elmass", k_elmass, nblocks(nel), kBlock, nel, h->R(BL_ELDENSITY),
            h->R(BL_ELVOLUME), h->R(BL_CNWT), h->R(BL_ELMASS), h->R(BL_CNMASS));
  if (mix_initmass(h)) return BL_FAILURE;
  return bl_hydro_geteos(h);
}

int bl_hydro_init(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nel2 = h->sz.nel2;
  BL_LAUNCH(h, "fill", k_fill_r, nblocks(nel2), kBlock, h->R(BL_ELVISC), (int64_t)nel2, 0.0);
  BL_LAUNCH(h, "fill", k_fill_r, nblocks(4 * (int64_t)nel2), kBlock, h->R(BL_CNVISCX), 4 * (int64_t)nel2, 0.0);
  BL_LAUNCH(h, "fill", k_fill_r, nblocks(4 * (int64_t)nel2), kBlock, h->R(BL_CNVISCY), 4 * (int64_t)nel2, 0.0);
  if (h->cfg.zsp)
    BL_LAUNCH(h, "spmassinit", k_spmassinit, nblocks(h->sz.nel), kBlock, h->sz.nel,
              h->R(BL_ELDENSITY), h->R(BL_CNX), h->R(BL_CNY), h->R(BL_SPMASS));
  if (mix_viscinit(h)) return BL_FAILURE;
  BL_HIP(hipStreamSynchronize(h->stream));
  return BL_SUCCESS;
}

// ----------------------------------------------------------- Lagrangian step
// hydro_dr_getdt with the partials at (part, parti), each holding 3 * nb
// slots for nb = nblocks(nel), and the result at dres
static int getdt_with(bl_handle_t* h, double* part, int32_t* parti, DtResult* dres,
                      bl_dt_t* cfl, bl_dt_t* div, bl_error_t* err) {
  int nel = h->sz.nel;
  unsigned nb = nblocks(nel);
  const bl_config_t& c = h->cfg;
  BL_LAUNCH(h, "getdt_a", k_getdt_a, nb, kBlock, nel, h->I(BL_IELND), h->R(BL_NDU),
            h->R(BL_NDV), h->R(BL_CNX), h->R(BL_CNY), h->R(BL_A1), h->R(BL_A3), h->R(BL_B1),
            h->R(BL_B3), h->R(BL_ELVOLUME), h->R(BL_RSCRATCH21), h->R(BL_RSCRATCH22),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24), h->R(BL_RSCRATCH25), h->R(BL_RSCRATCH26),
            part, parti);
  BL_LAUNCH(h, "getdt_b", k_getdt_b, nb, kBlock, nel, c.zerocut, c.cvisc1, c.cvisc2, c.zcut,
            h->I(BL_INDTYPE), h->I(BL_IELEL), h->I(BL_IELND), h->I(BL_IELFC), h->I(BL_IELREG),
            h->d_zdtnotreg, h->d_zmidlength, h->R(BL_ELDENSITY), h->R(BL_ELCS2),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24), h->R(BL_RSCRATCH25), h->R(BL_RSCRATCH26),
            h->R(BL_CNX), h->R(BL_CNY), h->R(BL_RSCRATCH21), h->R(BL_RSCRATCH22),
            h->R(BL_RSCRATCH27), h->R(BL_CNVISCX), h->R(BL_CNVISCY), h->R(BL_ELVISC),
            h->R(BL_RSCRATCH12));
  BL_LAUNCH(h, "getdt_c", k_getdt_c, nb, kBlock, nel, h->I(BL_IELREG), h->d_zdtnotreg,
            h->R(BL_ELVISC), h->R(BL_ELCS2), h->R(BL_RSCRATCH12), h->R(BL_RSCRATCH11),
            part, parti, (int)nb, (const int32_t*)h->I(BL_IELMAT), make_mix(h));
  BL_LAUNCH(h, "dt_final", k_dt_final, 1, kFinalBlock, (int)nb, (int)nb, part, parti, dres);
  BL_HIP(hipMemcpyAsync(h->h_dtres, dres, sizeof(DtResult), hipMemcpyDeviceToHost, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  const DtResult& r = *h->h_dtres;
  if (r.cfl_min < 0.0) {
    if (err) {
      err->ierr = BL_FAILURE; err->iout = BL_HALT_SINGLE;
      std::memset(err->serr, ' ', BL_SLEN);
      const char* m = "ERROR: negative CFL condition in hydro_dr_getdt";
      std::memcpy(err->serr, m, std::strlen(m));
    }
    set_error("negative CFL condition in hydro_dr_getdt");
    return BL_FAILURE;
  }
  if (cfl) { cfl->rdt = c.cfl_sf * sqrt(r.cfl_min); cfl->idt = r.cfl_idx; cfl->sdt = BL_DT_CFL; }
  if (div) { div->rdt = c.div_sf / r.div_max; div->idt = r.div_idx; div->sdt = BL_DT_DIV; }
  if (err) { err->ierr = BL_SUCCESS; err->iout = 0; std::memset(err->serr, ' ', BL_SLEN); }
  return BL_SUCCESS;
}

int bl_hydro_getdt(bl_handle_t* h, bl_dt_t* cfl, bl_dt_t* div, bl_error_t* err) {
  bl_touch(h);
  if (!h || !h->d_zdtnotreg) { set_error("bl_hydro_getdt: regions not set"); return BL_FAILURE; }
  return getdt_with(h, h->d_part, h->d_parti, h->d_dtres, cfl, div, err);
}

int bl_hydro_getdt_noalloc(bl_handle_t* h, bl_dt_t* cfl, bl_dt_t* div, bl_error_t* err) {
  bl_touch(h);
  if (!h || !h->d_zdtnotreg) {
    set_error("bl_hydro_getdt_noalloc: regions not set");
    return BL_FAILURE;
  }
  if (h->sz.nel > BL_NOALLOC_MAX_EL) {
    set_error("bl_hydro_getdt_noalloc: " + std::to_string(h->sz.nel) +
              " elements exceed the workspace maximum of " +
              std::to_string(BL_NOALLOC_MAX_EL) + " (use bl_hydro_getdt)");
    return BL_FAILURE;
  }
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != h->device) {
    set_error("bl_hydro_getdt_noalloc: the handle's device is not the current device");
    return BL_FAILURE;
  }
  unsigned char* ws = noalloc_scratch(h->device);
  if (!ws) return BL_FAILURE;
  return getdt_with(h, reinterpret_cast<double*>(ws + kNaPart),
                    reinterpret_cast<int32_t*>(ws + kNaParti), reinterpret_cast<DtResult*>(ws),
                    cfl, div, err);
}

int bl_hydro_setpredictor(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nel = h->sz.nel;
  BL_LAUNCH(h, "copy", k_copy_r, nblocks(nel), kBlock, h->R(BL_ELENERGY), h->R(BL_RSCRATCH11), nel);
  return BL_SUCCESS;
}

int bl_hydro_setcorrector(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nel = h->sz.nel;
  BL_LAUNCH(h, "copy", k_copy_r, nblocks(nel), kBlock, h->R(BL_RSCRATCH11), h->R(BL_ELENERGY), nel);
  return BL_SUCCESS;
}

int bl_hydro_getforce(bl_handle_t* h, const bl_timestep_t* ts) {
  bl_touch(h);
  if (!h || !ts) return BL_FAILURE;
  int nel = h->sz.nel;
  int zhg = (ts->zcorrector && h->cfg.zhg) ? 1 : 0;
  BL_LAUNCH(h, "getforce", k_getforce, nblocks(nel), kBlock, nel, h->R(BL_ELPRESSURE),
            h->R(BL_A1), h->R(BL_A3), h->R(BL_B1), h->R(BL_B3), h->R(BL_CNVISCX),
            h->R(BL_CNVISCY), h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24), h->cfg.zsp, zhg, ts->dt,
            h->I(BL_IELREG), h->R(BL_ELDENSITY), h->R(BL_ELCS2), h->R(BL_CNX), h->R(BL_CNY),
            h->R(BL_SPMASS), h->R(BL_ELVOLUME), h->R(BL_RSCRATCH21), h->R(BL_RSCRATCH22),
            h->d_kappareg, h->d_pmeritreg);
  return mix_force(h);
}

int bl_hydro_getstate(bl_handle_t* h, const bl_timestep_t* ts, bl_error_t* err) {
  bl_touch(h);
  if (!h || !ts) return BL_FAILURE;
  int nnd = h->sz.nnd;
  BL_LAUNCH(h, "getvertex", k_getvertex, nblocks(nnd), kBlock, nnd, 0.5 * ts->dt, h->R(BL_NDX),
            h->R(BL_NDY), h->R(BL_NDU), h->R(BL_NDV));
  if (bl_geometry_getgeometry(h, err) != BL_SUCCESS) return BL_FAILURE;
  return bl_density_getdensity(h);
}

int bl_hydro_getenergy(bl_handle_t* h, const bl_timestep_t* ts) {
  bl_touch(h);
  if (!h || !ts) return BL_FAILURE;
  int nel = h->sz.nel;
  BL_LAUNCH(h, "getenergy", k_getenergy, nblocks(nel), kBlock, nel, ts->dts, h->cfg.zerocut,
            h->R(BL_ELENERGY), h->R(BL_ELMASS), h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24),
            h->R(BL_RSCRATCH21), h->R(BL_RSCRATCH22));
  return mix_energy(h, ts->dts);
}

int bl_hydro_initacceleration(bl_handle_t* h) {
  bl_touch(h);
  // mixed cells: element corner forces from the components
  // (hydro_dr_getacceleration.f90:69-72); nproc=1: no exchange
  if (!h) return BL_FAILURE;
  return mix_average_force(h);
}

int bl_hydro_scatteracceleration(bl_handle_t* h) {
  bl_touch(h);
  if (!h || !h->maps_built) { set_error("bl_hydro_scatteracceleration: maps not built"); return BL_FAILURE; }
  int nnd = h->sz.nnd1;
  BL_LAUNCH(h, "scatteracc", k_scatteracc, nblocks(nnd), kBlock, nnd, h->d_hyd_ptr, h->d_hyd_idx,
            h->cfg.zerocut, h->R(BL_ELDENSITY), h->R(BL_CNWT), h->R(BL_CNMASS),
            h->R(BL_RSCRATCH23), h->R(BL_RSCRATCH24), h->R(BL_RSCRATCH12), h->R(BL_RSCRATCH13),
            h->R(BL_RSCRATCH14), h->R(BL_RSCRATCH15));
  return BL_SUCCESS;
}

int bl_hydro_getacceleration(bl_handle_t* h) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  int nnd = h->sz.nnd;
  BL_LAUNCH(h, "getacc", k_getacc, nblocks(nnd), kBlock, nnd, h->cfg.dencut, h->cfg.zerocut,
            h->cfg.accut * h->cfg.accut, h->R(BL_RSCRATCH12), h->R(BL_RSCRATCH13),
            h->R(BL_RSCRATCH14), h->R(BL_RSCRATCH15), h->I(BL_INDTYPE));
  return BL_SUCCESS;
}

int bl_hydro_applyacceleration(bl_handle_t* h, const bl_timestep_t* ts) {
  bl_touch(h);
  if (!h || !ts) return BL_FAILURE;
  int nnd = h->sz.nnd, nel = h->sz.nel;
  BL_LAUNCH(h, "applyacc", k_applyacc, nblocks(nnd), kBlock, nnd, ts->dt, h->R(BL_RSCRATCH14),
            h->R(BL_RSCRATCH15), h->R(BL_NDU), h->R(BL_NDV));
  BL_LAUNCH(h, "cngather", k_cngather2, nblocks(nel), kBlock, nel, h->I(BL_IELND),
            h->R(BL_RSCRATCH14), h->R(BL_RSCRATCH15), h->R(BL_RSCRATCH21), h->R(BL_RSCRATCH22));
  return BL_SUCCESS;
}

// ---------------------------------------------------------------- timing
int bl_timing_enable(bl_handle_t* h, int32_t on) {
  if (!h) return BL_FAILURE;
  if (on && h->ev_pool.empty()) {
    const int npairs = 8192;
    h->ev_pool.resize(2 * npairs);
    h->ev_class.resize(npairs);
    for (auto& ev : h->ev_pool) BL_HIP(hipEventCreate(&ev));
  }
  if (!on) timing_flush(h);
  h->timing = on != 0;
  return BL_SUCCESS;
}

int bl_timing_reset(bl_handle_t* h) {
  if (!h) return BL_FAILURE;
  timing_flush(h);
  for (auto& c : h->tclass) { c.total_ms = 0.0; c.launches = 0; }
  return BL_SUCCESS;
}

int bl_timing_query(bl_handle_t* h, const char* name, double* avg_ms, int64_t* launches) {
  if (!h || !name) return BL_FAILURE;
  timing_flush(h);
  for (auto& c : h->tclass)
    if (c.name == name) {
      if (avg_ms) *avg_ms = c.launches ? c.total_ms / c.launches : 0.0;
      if (launches) *launches = c.launches;
      return BL_SUCCESS;
    }
  if (avg_ms) *avg_ms = 0.0;
  if (launches) *launches = 0;
  return BL_SUCCESS;
}

int bl_timing_names(bl_handle_t* h, char* buf, int32_t buflen) {
  if (!h || !buf || buflen <= 0) return BL_FAILURE;
  std::string s;
  for (auto& c : h->tclass) { if (!s.empty()) s += ","; s += c.name; }
  std::strncpy(buf, s.c_str(), buflen - 1);
  buf[buflen - 1] = 0;
  return BL_SUCCESS;
}

}  // extern "C"
