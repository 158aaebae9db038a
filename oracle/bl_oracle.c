/*
 * bl_oracle.c -- CPU restatement of BookLeaf's Lagrangian step and ALE remap.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
 * product.  Build: gcc -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile).
 *
 * Each routine restates the cited reference routine statement by statement,
 * keeping Fortran's left-to-right association so the results are bit-equal
 * to the reference compiled without FMA.  Index arrays hold the reference's
 * 1-based values; CN(a,j,e) addresses the Fortran element a(j,e) of an
 * (NCORN,n) array and EL(a,e) the element a(e).
 *
 * Parity pinning: tests/test_oracle_ref.py compares every driver below with
 * the reference's own Fortran (oracle/_ref, built by oracle/build_ref.py) on
 * the same inputs, bit for bit; tests/test_oracle_invariants.py checks the
 * closed forms (sum of corner forces, sum of corner weights, ideal-gas EOS,
 * uniform-state invariance, CFL of a uniform mesh, analytic Sod L1).
 */
#include "bl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CN(a, j, e) (a)[(int64_t)((e)-1) * 4 + ((j)-1)]
#define EL(a, e) (a)[(e)-1]

static inline double fmax_f(double a, double b) { return (a > b) ? a : b; }
static inline double fmin_f(double a, double b) { return (a < b) ? a : b; }
/* Fortran SIGN(a,b): |a| with the sign of b */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
/* Fortran MOD for non-negative operands and MODULO */
static inline int fmod_i(int a, int p) { return a % p; }
static inline int fmodulo_i(int a, int p) {
  int r = a % p;
  return (r < 0) ? r + p : r;
}

struct orc_state {
  bl_sizes_t sz;
  bl_config_t cfg;
  /* mesh-motion option of the ALE step (bl_set_relaxation) */
  int32_t relax_type, relax_nsweep;
  double relax_om, relax_tol;
  int32_t* ring; /* (8, nnd) 0-based stencil nodes, -1: node not relaxed */
  int32_t relax_plant; /* tests only: a deliberate mistake in the stencil */
  int32_t nreg;
  double* kappareg;
  double* pmeritreg;
  int32_t* zdtnotreg;
  int32_t* zmidlength;
  int32_t nmat;
  int32_t* eos_itype; /* 0..nmat */
  double* eos_param;  /* (6, 0:nmat) */
  void* f[BL_NFIELD + 1];
  int32_t kind[BL_NFIELD + 1];
  int64_t count[BL_NFIELD + 1];
};

#define R(id) ((double*)s->f[id])
#define I(id) ((int32_t*)s->f[id])
#define Z(id) ((uint8_t*)s->f[id])

/* shapes of data.f90:169-430: returns kind and element count */
static void field_shape(const bl_sizes_t* z, int id, int32_t* kind,
                        int64_t* count) {
  int64_t nel = z->nel2, nnd = z->nnd2, nsz = z->nsz, mcp = z->mcp,
          mmx = z->mmx;
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

orc_state_t* orc_create(const bl_sizes_t* sizes) {
  orc_state_t* s = (orc_state_t*)calloc(1, sizeof(orc_state_t));
  s->sz = *sizes;
  s->sz.nsz = (s->sz.nel2 > s->sz.nnd2) ? s->sz.nel2 : s->sz.nnd2;
  for (int id = 1; id <= BL_NFIELD; ++id) {
    field_shape(&s->sz, id, &s->kind[id], &s->count[id]);
    int64_t n = s->count[id] > 0 ? s->count[id] : 1;
    if (s->kind[id] == BL_KIND_R64) {
      double* p = (double*)malloc(sizeof(double) * n);
      for (int64_t i = 0; i < n; ++i) p[i] = -2.0e12; /* RINIT */
      s->f[id] = p;
    } else if (s->kind[id] == BL_KIND_I32) {
      int32_t* p = (int32_t*)malloc(sizeof(int32_t) * n);
      for (int64_t i = 0; i < n; ++i) p[i] = -2000000000; /* IINIT */
      s->f[id] = p;
    } else {
      s->f[id] = calloc(n, 1); /* ZINIT=.FALSE. */
    }
  }
  return s;
}

void orc_destroy(orc_state_t* s) {
  if (!s) return;
  for (int id = 1; id <= BL_NFIELD; ++id) free(s->f[id]);
  free(s->kappareg); free(s->pmeritreg); free(s->zdtnotreg);
  free(s->zmidlength); free(s->eos_itype); free(s->eos_param);
  free(s->ring);
  free(s);
}

void orc_set_config(orc_state_t* s, const bl_config_t* cfg) { s->cfg = *cfg; }

void orc_set_regions(orc_state_t* s, int32_t nreg, const double* kappareg,
                     const double* pmeritreg, const int32_t* zdtnotreg,
                     const int32_t* zmidlength) {
  free(s->kappareg); free(s->pmeritreg); free(s->zdtnotreg);
  free(s->zmidlength);
  s->nreg = nreg;
  s->kappareg = (double*)malloc(sizeof(double) * nreg);
  s->pmeritreg = (double*)malloc(sizeof(double) * nreg);
  s->zdtnotreg = (int32_t*)malloc(sizeof(int32_t) * nreg);
  s->zmidlength = (int32_t*)malloc(sizeof(int32_t) * nreg);
  memcpy(s->kappareg, kappareg, sizeof(double) * nreg);
  memcpy(s->pmeritreg, pmeritreg, sizeof(double) * nreg);
  memcpy(s->zdtnotreg, zdtnotreg, sizeof(int32_t) * nreg);
  memcpy(s->zmidlength, zmidlength, sizeof(int32_t) * nreg);
}

void orc_set_eos(orc_state_t* s, int32_t nmat, const int32_t* itype,
                 const double* param) {
  free(s->eos_itype); free(s->eos_param);
  s->nmat = nmat;
  s->eos_itype = (int32_t*)malloc(sizeof(int32_t) * (nmat + 1));
  s->eos_param = (double*)malloc(sizeof(double) * BL_NEOSPARAM * (nmat + 1));
  memcpy(s->eos_itype, itype, sizeof(int32_t) * (nmat + 1));
  memcpy(s->eos_param, param, sizeof(double) * BL_NEOSPARAM * (nmat + 1));
}

void* orc_field(orc_state_t* s, int32_t id, int32_t* kind, int64_t* count) {
  if (id < 1 || id > BL_NFIELD) return NULL;
  if (kind) *kind = s->kind[id];
  if (count) *count = s->count[id];
  return s->f[id];
}

void orc_get_sizes(orc_state_t* s, bl_sizes_t* out) { *out = s->sz; }

static void set_err(bl_error_t* err, int ierr, int iout, const char* msg) {
  if (!err) return;
  err->ierr = ierr;
  err->iout = iout;
  memset(err->serr, ' ', BL_SLEN);
  size_t n = strlen(msg);
  memcpy(err->serr, msg, n < BL_SLEN ? n : BL_SLEN);
}

/* ===================================================================== */
/* utilities                                                              */
/* ===================================================================== */

/* utils_kn_cngather, utils_kn_gather.f90:35-51 */
static void k_cngather(int n1, const int32_t* indx, const double* aa,
                       double* bb) {
  for (int ii = 1; ii <= n1; ++ii)
    for (int jj = 1; jj <= 4; ++jj) CN(bb, jj, ii) = EL(aa, CN(indx, jj, ii));
}

/* utils_kn_copyr1, utils_kn_copy.f90:33-46 */
static void k_copy(int n, const double* right, double* left) {
  for (int ii = 0; ii < n; ++ii) left[ii] = right[ii];
}

/* utils_kn_divide / utils_kn_sum, utils_kn_math.f90:28-56 */
static void k_divide(int n, const double* num, const double* den,
                     double* ans) {
  for (int ii = 0; ii < n; ++ii) ans[ii] = num[ii] / den[ii];
}
static void k_sum(int n, const double* right, double* ans) {
  for (int ii = 0; ii < n; ++ii) ans[ii] = ans[ii] + right[ii];
}

/* utils_kn_bc, utils_kn_bc.f90:28-58 */
static void k_bc(int nnd, double rcut, const int32_t* indtype, double* ndu,
                 double* ndv) {
  double w1 = rcut * rcut;
  for (int ind = 1; ind <= nnd; ++ind) {
    switch (EL(indtype, ind)) {
      case -1: EL(ndu, ind) = 0.0; break;
      case -2: EL(ndv, ind) = 0.0; break;
      case -3: EL(ndu, ind) = 0.0; EL(ndv, ind) = 0.0; break;
      default: break;
    }
    double w2 = EL(ndu, ind) * EL(ndu, ind) + EL(ndv, ind) * EL(ndv, ind);
    if (w2 < w1) {
      EL(ndu, ind) = 0.0;
      EL(ndv, ind) = 0.0;
    }
  }
}

/* ===================================================================== */
/* geometry                                                               */
/* ===================================================================== */

/* geometry_kn_getiso, geometry_kn_volume.f90:30-71 */
static void k_getiso(int nel, const double* cnx, const double* cny,
                     double* a1, double* a2, double* a3, double* b1,
                     double* b2, double* b3, double* cnwt) {
  const double ONEBYNINE = 1.0 / 9.0;
  for (int iel = 1; iel <= nel; ++iel) {
    double x1 = CN(cnx, 1, iel), x2 = CN(cnx, 2, iel), x3 = CN(cnx, 3, iel),
           x4 = CN(cnx, 4, iel);
    double y1 = CN(cny, 1, iel), y2 = CN(cny, 2, iel), y3 = CN(cny, 3, iel),
           y4 = CN(cny, 4, iel);
    double A1 = 0.25 * (-x1 + x2 + x3 - x4);
    double A2 = 0.25 * (x1 - x2 + x3 - x4);
    double A3 = 0.25 * (-x1 - x2 + x3 + x4);
    double B1 = 0.25 * (-y1 + y2 + y3 - y4);
    double B2 = 0.25 * (y1 - y2 + y3 - y4);
    double B3 = 0.25 * (-y1 - y2 + y3 + y4);
    EL(a1, iel) = A1; EL(a2, iel) = A2; EL(a3, iel) = A3;
    EL(b1, iel) = B1; EL(b2, iel) = B2; EL(b3, iel) = B3;
    CN(cnwt, 1, iel) = ONEBYNINE * ((3.0 * B3 - B2) * (3.0 * A1 - A2) -
                                    (3.0 * A3 - A2) * (3.0 * B1 - B2));
    CN(cnwt, 2, iel) = ONEBYNINE * ((3.0 * B3 + B2) * (3.0 * A1 - A2) -
                                    (3.0 * A3 + A2) * (3.0 * B1 - B2));
    CN(cnwt, 3, iel) = ONEBYNINE * ((3.0 * B3 + B2) * (3.0 * A1 + A2) -
                                    (3.0 * A3 + A2) * (3.0 * B1 + B2));
    CN(cnwt, 4, iel) = ONEBYNINE * ((3.0 * B3 - B2) * (3.0 * A1 + A2) -
                                    (3.0 * A3 - A2) * (3.0 * B1 + B2));
  }
}

/* geometry_kn_getvolume, geometry_kn_volume.f90:73-86 */
static void k_getvolume(int n, const double* a1, const double* a3,
                        const double* b1, const double* b3, double* vol) {
  for (int ii = 0; ii < n; ++ii) vol[ii] = 4.0 * (a1[ii] * b3[ii] - a3[ii] * b1[ii]);
}

/* geometry_kn_checkvolume, geometry_kn_volume.f90:88-99: last offender */
static int k_checkvolume(int nel, double val, const double* vol) {
  int ierr = BL_SUCCESS;
  for (int iel = 1; iel <= nel; ++iel)
    if (EL(vol, iel) < val) ierr = -iel;
  return ierr;
}

/* geometry_kn_getfluxvolume, geometry_kn_volume.f90:101-140 */
static void k_getfluxvolume(int nel, double cut, const int32_t* ielnd,
                            const double* ndx0, const double* ndy0,
                            const double* ndx1, const double* ndy1,
                            double* fcdv) {
  for (int iel = 1; iel <= nel; ++iel) {
    for (int jj = 1; jj <= 4; ++jj) {
      int jp = fmod_i(jj, 4) + 1;
      int n1 = CN(ielnd, jj, iel), n2 = CN(ielnd, jp, iel);
      double x1 = EL(ndx0, n1), x2 = EL(ndx0, n2);
      double y1 = EL(ndy0, n1), y2 = EL(ndy0, n2);
      double x3 = EL(ndx1, n2), x4 = EL(ndx1, n1);
      double y3 = EL(ndy1, n2), y4 = EL(ndy1, n1);
      double a1 = (-x1 - x4) + (x3 + x2);
      double a3 = (-x1 + x4) + (x3 - x2);
      double b1 = (-y1 - y4) + (y3 + y2);
      double b3 = (-y1 + y4) + (y3 - y2);
      double v = 0.25 * (a1 * b3 - a3 * b1);
      if (v < cut) v = 0.0;
      CN(fcdv, jj, iel) = v;
    }
  }
}

/* geometry_kn_length.f90:90-133: denom, distpp, distpl */
static inline double g_denom(double x1, double y1, double x2, double y2) {
  double w1 = y1 - y2, w2 = x1 - x2;
  return w1 * w1 + w2 * w2;
}
static inline double g_distpp(double x3, double y3, double x4, double y4,
                              double x1, double y1) {
  double w1 = 0.5 * (x3 + x4) - x1, w2 = 0.5 * (y3 + y4) - y1;
  return w1 * w1 + w2 * w2;
}
static inline double g_distpl(double x3, double y3, double x4, double y4,
                              double x1, double y1, double x2, double y2) {
  double r = 0.5 * (y1 - y2) * (x3 + x4) + 0.5 * (y3 + y4) * (x2 - x1) +
             y2 * x1 - y1 * x2;
  return r * r;
}
/* geometry_kn_dln, geometry_kn_length.f90:57-88 */
static void g_dln(const double* x, const double* y, double zcut,
                  double* res) {
  double w1;
  w1 = g_denom(x[2], y[2], x[3], y[3]);
  res[0] = (w1 < zcut) ? g_distpp(x[0], y[0], x[1], y[1], x[2], y[2])
                       : g_distpl(x[0], y[0], x[1], y[1], x[2], y[2], x[3], y[3]) / w1;
  w1 = g_denom(x[3], y[3], x[0], y[0]);
  res[1] = (w1 < zcut) ? g_distpp(x[1], y[1], x[2], y[2], x[3], y[3])
                       : g_distpl(x[1], y[1], x[2], y[2], x[3], y[3], x[0], y[0]) / w1;
  w1 = g_denom(x[0], y[0], x[1], y[1]);
  res[2] = (w1 < zcut) ? g_distpp(x[2], y[2], x[3], y[3], x[0], y[0])
                       : g_distpl(x[2], y[2], x[3], y[3], x[0], y[0], x[1], y[1]) / w1;
  w1 = g_denom(x[1], y[1], x[2], y[2]);
  res[3] = (w1 < zcut) ? g_distpp(x[3], y[3], x[0], y[0], x[1], y[1])
                       : g_distpl(x[3], y[3], x[0], y[0], x[1], y[1], x[2], y[2]) / w1;
}
/* geometry_kn_dlm, geometry_kn_length.f90:30-55 */
static void g_dlm(const double* x, const double* y, double* res) {
  double x1 = x[0] + x[1], x2 = x[2] + x[3];
  double y1 = y[0] + y[1], y2 = y[2] + y[3];
  x1 = 0.5 * (x1 - x2);
  y1 = 0.5 * (y1 - y2);
  res[0] = x1 * x1 + y1 * y1;
  x1 = x[2] + x[1]; x2 = x[0] + x[3];
  y1 = y[2] + y[1]; y2 = y[0] + y[3];
  x1 = 0.5 * (x1 - x2);
  y1 = 0.5 * (y1 - y2);
  res[1] = x1 * x1 + y1 * y1;
  res[2] = res[0];
  res[3] = res[1];
}

/* geometry_kn_getvertex, geometry_kn_position.f90:28-43 */
static void k_getvertex(int nnd, double dt, double* ndx, double* ndy,
                        const double* ndu, const double* ndv) {
  for (int ind = 0; ind < nnd; ++ind) {
    ndx[ind] = ndx[ind] + dt * ndu[ind];
    ndy[ind] = ndy[ind] + dt * ndv[ind];
  }
}

/* ===================================================================== */
/* EOS: eos_kn_forms.f90:28-106                                           */
/* ===================================================================== */
static double eos_pressure(const orc_state_t* s, int mat, double density,
                           double energy) {
  int imat = mat > 0 ? mat : 0;
  const double* p = s->eos_param + (int64_t)BL_NEOSPARAM * imat;
  double pressure, t1, t2, t3, t4, t5;
  switch (s->eos_itype[imat]) {
    case 0: pressure = p[0]; break;
    case 1: pressure = energy * density * (p[0] - 1.0); break;
    case 2:
      t1 = density / p[2];
      pressure = p[0] * (pow(t1, p[1]) - 1.0);
      pressure = fmax_f(pressure, p[3]);
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
  if (fabs(pressure) < s->cfg.pcut) pressure = 0.0;
  return pressure;
}
static double eos_cs2(const orc_state_t* s, int mat, double density,
                      double energy) {
  int imat = mat > 0 ? mat : 0;
  const double* p = s->eos_param + (int64_t)BL_NEOSPARAM * imat;
  double cs2, t1, t2, t3, t4, t5;
  switch (s->eos_itype[imat]) {
    case 0: cs2 = s->cfg.ccut; break;
    case 1: cs2 = p[0] * (p[0] - 1.0) * energy; break;
    case 2:
      t1 = density / p[2];
      t2 = p[1] - 1.0;
      cs2 = (p[0] * p[1]) / p[2];
      cs2 = cs2 * pow(t1, t2);
      break;
    case 3:
      t1 = p[5] / density;
      t2 = eos_pressure(s, imat, density, energy);
      t3 = p[3] * t1;
      t4 = p[0] / p[3] + p[0] * t1 - t3 * t1;
      t4 = t4 * p[1] * exp(-t3);
      t3 = p[4] * t1;
      t5 = p[0] / p[4] + p[0] * t1 - t3 * t1;
      t5 = t5 * p[2] * exp(-t3);
      cs2 = p[0] * t2 / density + p[0] * energy - t4 - t5;
      break;
    default: cs2 = s->cfg.ccut;
  }
  return fmax_f(cs2, s->cfg.ccut);
}
/* eos_kn_geteos, eos_kn_geteos.f90:28-52 */
static void k_geteos(const orc_state_t* s, int n, const int32_t* imat,
                     const double* den, const double* en, double* pre,
                     double* cs2) {
  for (int ii = 0; ii < n; ++ii) {
    pre[ii] = eos_pressure(s, imat[ii], den[ii], en[ii]);
    cs2[ii] = eos_cs2(s, imat[ii], den[ii], en[ii]);
  }
}

/* ===================================================================== */
/* hydro kernels                                                          */
/* ===================================================================== */

/* hydro_kn_initeviscosity, hydro_kn_edgeviscosity.f90:30-71 */
static void k_initeviscosity(int nel, const double* cnx, const double* cny,
                             const double* cnu, const double* cnv,
                             double* elvisc, double* dx, double* dy,
                             double* du, double* dv, double* edx,
                             double* edy) {
  for (int iel = 1; iel <= nel; ++iel) {
    EL(elvisc, iel) = 0.0;
    for (int j = 1; j <= 4; ++j) {
      CN(edx, j, iel) = 0.0;
      CN(edy, j, iel) = 0.0;
    }
    for (int j = 1; j <= 4; ++j) {
      int jp = (j % 4) + 1;
      CN(du, j, iel) = CN(cnu, jp, iel) - CN(cnu, j, iel);
      CN(dv, j, iel) = CN(cnv, jp, iel) - CN(cnv, j, iel);
      CN(dx, j, iel) = CN(cnx, jp, iel) - CN(cnx, j, iel);
      CN(dy, j, iel) = CN(cny, jp, iel) - CN(cny, j, iel);
    }
  }
}

/* hydro_kn_limiteviscosity, hydro_kn_edgeviscosity.f90:73-217 */
static void k_limiteviscosity(int nel, double zerocut, double cvisc1,
                              double cvisc2, const int32_t* indtype,
                              const int32_t* ielel, const int32_t* ielnd,
                              const int32_t* ielfc, const double* den,
                              const double* cs2, const double* du,
                              const double* dv, const double* dx,
                              const double* dy, double* scratch,
                              double* edx, double* edy, double* elvisc) {
  for (int iside = 1; iside <= 2; ++iside) {
    int is1 = fmod_i(iside + 2, 4) + 1;
    int is2 = iside + 1;
    for (int iel = 1; iel <= nel; ++iel) {
      int in1 = CN(ielel, iside, iel);
      int in2 = CN(ielel, iside + 2, iel);
      double w1, w2, w3, w4, den_, uhat, vhat, xhat, yhat;
      int ins;
      /* edge 1 */
      w1 = CN(du, is1, iel); w2 = CN(dv, is1, iel);
      w3 = CN(dx, is1, iel); w4 = CN(dy, is1, iel);
      den_ = sqrt(w1 * w1 + w2 * w2);
      den_ = 1.0 / fmax_f(den_, zerocut);
      uhat = w1 * den_; vhat = w2 * den_;
      den_ = sqrt(w3 * w3 + w4 * w4);
      den_ = 1.0 / fmax_f(den_, zerocut);
      xhat = w3 * den_; yhat = w4 * den_;
      den_ = w3 * xhat + w4 * yhat;
      w1 = (w1 * uhat + w2 * vhat) / fsign(fmax_f(fabs(den_), zerocut), den_);
      w1 = 1.0 / fsign(fmax_f(fabs(w1), zerocut), w1);
      ins = CN(ielfc, iside, iel);
      ins = fmod_i(ins, 4) + 1;
      den_ = CN(dx, ins, in1) * xhat + CN(dy, ins, in1) * yhat;
      w2 = (CN(du, ins, in1) * uhat + CN(dv, ins, in1) * vhat) /
           fsign(fmax_f(fabs(den_), zerocut), den_);
      CN(scratch, 1, iel) = w2 * w1;
      ins = CN(ielfc, iside + 2, iel);
      ins = fmod_i(ins + 2, 4) + 1;
      den_ = CN(dx, ins, in2) * xhat + CN(dy, ins, in2) * yhat;
      w3 = (CN(du, ins, in2) * uhat + CN(dv, ins, in2) * vhat) /
           fsign(fmax_f(fabs(den_), zerocut), den_);
      CN(scratch, 2, iel) = w3 * w1;
      /* edge 2 */
      w1 = CN(du, is2, iel); w2 = CN(dv, is2, iel);
      w3 = CN(dx, is2, iel); w4 = CN(dy, is2, iel);
      den_ = sqrt(w1 * w1 + w2 * w2);
      den_ = 1.0 / fmax_f(den_, zerocut);
      uhat = w1 * den_; vhat = w2 * den_;
      den_ = sqrt(w3 * w3 + w4 * w4);
      den_ = 1.0 / fmax_f(den_, zerocut);
      xhat = w3 * den_; yhat = w4 * den_;
      den_ = w3 * xhat + w4 * yhat;
      w1 = (w1 * uhat + w2 * vhat) / fsign(fmax_f(fabs(den_), zerocut), den_);
      w1 = 1.0 / fsign(fmax_f(fabs(w1), zerocut), w1);
      ins = CN(ielfc, iside, iel);
      ins = fmod_i(ins + 2, 4) + 1;
      den_ = CN(dx, ins, in1) * xhat + CN(dy, ins, in1) * yhat;
      w2 = (CN(du, ins, in1) * uhat + CN(dv, ins, in1) * vhat) /
           fsign(fmax_f(fabs(den_), zerocut), den_);
      CN(scratch, 3, iel) = w2 * w1;
      ins = CN(ielfc, iside + 2, iel);
      ins = fmod_i(ins, 4) + 1;
      den_ = CN(dx, ins, in2) * xhat + CN(dy, ins, in2) * yhat;
      w3 = (CN(du, ins, in2) * uhat + CN(dv, ins, in2) * vhat) /
           fsign(fmax_f(fabs(den_), zerocut), den_);
      CN(scratch, 4, iel) = w3 * w1;
    }
    /* BC */
    int insb = iside + 2;
    for (int iel = 1; iel <= nel; ++iel) {
      int in1 = CN(ielel, iside, iel);
      int in2 = CN(ielel, insb, iel);
      if (in1 == iel) {
        int ic1 = CN(ielnd, iside, iel);
        int ic2 = CN(ielnd, fmod_i(iside, 4) + 1, iel);
        if ((EL(indtype, ic1) < 0 && EL(indtype, ic2) < 0) && in2 != iel) {
          CN(scratch, 1, iel) = 1.0; CN(scratch, 3, iel) = 1.0;
        } else {
          CN(scratch, 1, iel) = 0.0; CN(scratch, 3, iel) = 0.0;
        }
      }
      if (in2 == iel) {
        int ic1 = CN(ielnd, insb, iel);
        int ic2 = CN(ielnd, fmod_i(insb, 4) + 1, iel);
        if ((EL(indtype, ic1) < 0 && EL(indtype, ic2) < 0) && in1 != iel) {
          CN(scratch, 2, iel) = 1.0; CN(scratch, 4, iel) = 1.0;
        } else {
          CN(scratch, 2, iel) = 0.0; CN(scratch, 4, iel) = 0.0;
        }
      }
    }
    /* apply limiter */
    for (int iel = 1; iel <= nel; ++iel) {
      int ins = CN(ielel, is1, iel);
      double w1 = cvisc1 * sqrt(0.5 * (EL(cs2, iel) + EL(cs2, ins)));
      double w2 = CN(scratch, 1, iel);
      double w3 = CN(scratch, 2, iel);
      w2 = fmin_f(fmin_f(fmin_f(0.5 * (w2 + w3), 2.0 * w2), 2.0 * w3), 1.0);
      w2 = fmax_f(0.0, w2);
      w3 = CN(du, is1, iel);
      double w4 = CN(dv, is1, iel);
      w4 = sqrt(w3 * w3 + w4 * w4);
      w3 = 0.5 * (1.0 - w2) * (EL(den, iel) + EL(den, ins)) * (w1 + cvisc2 * w4);
      CN(edx, is1, iel) = w3;
      CN(edy, is1, iel) = w3;
      w4 = 2.0 * w4 / fmax_f(EL(den, iel), zerocut);
      EL(elvisc, iel) = fmax_f(EL(elvisc, iel), w3 * w4);
      ins = CN(ielel, is2, iel);
      w1 = cvisc1 * sqrt(0.5 * (EL(cs2, iel) + EL(cs2, ins)));
      w2 = CN(scratch, 3, iel);
      w3 = CN(scratch, 4, iel);
      w2 = fmin_f(fmin_f(fmin_f(0.5 * (w2 + w3), 2.0 * w2), 2.0 * w3), 1.0);
      w2 = fmax_f(0.0, w2);
      w3 = CN(du, is2, iel);
      w4 = CN(dv, is2, iel);
      w4 = sqrt(w3 * w3 + w4 * w4);
      w3 = 0.5 * (1.0 - w2) * (EL(den, iel) + EL(den, ins)) * (w1 + cvisc2 * w4);
      CN(edx, is2, iel) = w3;
      CN(edy, is2, iel) = w3;
      w4 = 2.0 * w4 / fmax_f(EL(den, iel), zerocut);
      EL(elvisc, iel) = fmax_f(EL(elvisc, iel), w3 * w4);
    }
  }
}

/* hydro_kn_geteviscosity, hydro_kn_edgeviscosity.f90:219-282 */
static void k_geteviscosity(int nel, double zerocut, const double* cnx,
                            const double* cny, const double* cnu,
                            const double* cnv, double* cnviscx,
                            double* cnviscy) {
  for (int iel = 1; iel <= nel; ++iel) {
    double edx[4], edy[4];
    for (int iside = 1; iside <= 4; ++iside) {
      int ins = fmod_i(iside, 4) + 1;
      double w1 = CN(cnx, iside, iel);
      double w2 = CN(cnx, ins, iel);
      double w3 = 0.5 * (w1 + w2);
      w1 = w2 - w1;
      w2 = 0.25 * (CN(cnx, 1, iel) + CN(cnx, 2, iel) + CN(cnx, 3, iel) + CN(cnx, 4, iel));
      double w4 = CN(cny, iside, iel);
      double w5 = CN(cny, ins, iel);
      double w6 = 0.5 * (w4 + w5);
      w4 = w5 - w4;
      w5 = 0.25 * (CN(cny, 1, iel) + CN(cny, 2, iel) + CN(cny, 3, iel) + CN(cny, 4, iel));
      double w7 = sqrt((w2 - w3) * (w2 - w3) + (w5 - w6) * (w5 - w6));
      double w8 = sqrt(w1 * w1 + w4 * w4);
      double den = 1.0 / w7;
      double xhat = (w5 - w6) * den;
      double yhat = (w3 - w2) * den;
      den = 1.0 / w8;
      w1 = w1 * den;
      w2 = w4 * den;
      w3 = xhat * w1 + yhat * w2;
      den = -fsign(1.0, w3) * w7;
      xhat = xhat * den;
      yhat = yhat * den;
      double uhat = CN(cnu, ins, iel) - CN(cnu, iside, iel);
      double vhat = CN(cnv, ins, iel) - CN(cnv, iside, iel);
      w5 = sqrt((uhat * uhat) + (vhat * vhat));
      w6 = uhat * xhat + vhat * yhat;
      den = w6 / fmax_f(w5, zerocut);
      edx[iside - 1] = CN(cnviscx, iside, iel) * uhat * den;
      edy[iside - 1] = CN(cnviscy, iside, iel) * vhat * den;
      if (w5 <= zerocut || w6 <= zerocut || w7 <= zerocut || w8 <= zerocut) {
        edx[iside - 1] = 0.0;
        edy[iside - 1] = 0.0;
      }
    }
    CN(cnviscx, 1, iel) = edx[0] - edx[3];
    CN(cnviscx, 2, iel) = edx[1] - edx[0];
    CN(cnviscx, 3, iel) = edx[2] - edx[1];
    CN(cnviscx, 4, iel) = edx[3] - edx[2];
    CN(cnviscy, 1, iel) = edy[0] - edy[3];
    CN(cnviscy, 2, iel) = edy[1] - edy[0];
    CN(cnviscy, 3, iel) = edy[2] - edy[1];
    CN(cnviscy, 4, iel) = edy[3] - edy[2];
  }
}

/* hydro_kn_getdtcfl, hydro_kn_getdt.f90:30-85 */
static int k_getdtcfl(const orc_state_t* s, int nel, double zcut,
                      double cfl_sf, const int32_t* ielreg, const double* cs2,
                      const double* cnx, const double* cny, double* rs11,
                      double* rs12, double* rdt, int* idt) {
  for (int iel = 1; iel <= nel; ++iel) {
    int ireg = EL(ielreg, iel);
    if (s->zdtnotreg[ireg - 1]) {
      EL(rs11, iel) = DBL_MAX;
      EL(rs12, iel) = DBL_MIN;
    } else {
      double x[4], y[4], res[4];
      for (int j = 0; j < 4; ++j) {
        x[j] = CN(cnx, j + 1, iel);
        y[j] = CN(cny, j + 1, iel);
      }
      if (s->zmidlength[ireg - 1]) g_dlm(x, y, res);
      else g_dln(x, y, zcut, res);
      double w1 = res[0];
      for (int j = 1; j < 4; ++j) w1 = fmin_f(w1, res[j]);
      EL(rs11, iel) = w1 / EL(cs2, iel);
      EL(rs12, iel) = w1;
    }
  }
  int ii = 1;
  for (int iel = 2; iel <= nel; ++iel)
    if (EL(rs11, iel) < EL(rs11, ii)) ii = iel;
  double w1 = EL(rs11, ii);
  if (w1 < 0.0) return BL_FAILURE;
  *rdt = cfl_sf * sqrt(w1);
  *idt = ii;
  return BL_SUCCESS;
}

/* hydro_kn_getdtdiv, hydro_kn_getdt.f90:87-122 */
static void k_getdtdiv(int nel, double div_sf, const double* a1,
                       const double* a3, const double* b1, const double* b3,
                       const double* vol, const double* cnu,
                       const double* cnv, double* rdt, int* idt) {
  double w2 = DBL_MIN;
  int ii = 0;
  for (int iel = 1; iel <= nel; ++iel) {
    double A1 = EL(a1, iel), A3 = EL(a3, iel), B1 = EL(b1, iel),
           B3 = EL(b3, iel);
    double w1 = CN(cnu, 1, iel) * (-B3 + B1) + CN(cnv, 1, iel) * (A3 - A1) +
                CN(cnu, 2, iel) * (B3 + B1) + CN(cnv, 2, iel) * (-A3 - A1) +
                CN(cnu, 3, iel) * (B3 - B1) + CN(cnv, 3, iel) * (-A3 + A1) +
                CN(cnu, 4, iel) * (-B3 - B1) + CN(cnv, 4, iel) * (A3 + A1);
    w1 = fabs(w1) / EL(vol, iel);
    if (w1 > w2) {
      w2 = w1;
      ii = iel;
    }
  }
  *rdt = div_sf / w2;
  *idt = ii;
}

/* hydro_kn_getforceP, hydro_kn_getforce.f90:30-53 */
static void k_getforceP(int nel, const double* p, const double* a1,
                        const double* a3, const double* b1, const double* b3,
                        double* cnfx, double* cnfy) {
  for (int iel = 1; iel <= nel; ++iel) {
    double w1 = EL(p, iel);
    double A1 = EL(a1, iel), A3 = EL(a3, iel), B1 = EL(b1, iel),
           B3 = EL(b3, iel);
    CN(cnfx, 1, iel) = w1 * (-B3 + B1);
    CN(cnfx, 2, iel) = w1 * (B3 + B1);
    CN(cnfx, 3, iel) = w1 * (B3 - B1);
    CN(cnfx, 4, iel) = w1 * (-B3 - B1);
    CN(cnfy, 1, iel) = w1 * (A3 - A1);
    CN(cnfy, 2, iel) = w1 * (-A3 - A1);
    CN(cnfy, 3, iel) = w1 * (-A3 + A1);
    CN(cnfy, 4, iel) = w1 * (A3 + A1);
  }
}

/* hydro_kn_getforceQ, hydro_kn_getforce.f90:55-71 */
static void k_getforceQ(int nel, const double* qx, const double* qy,
                        double* cnfx, double* cnfy) {
  for (int64_t i = 0; i < 4 * (int64_t)nel; ++i) {
    cnfx[i] = cnfx[i] + qx[i];
    cnfy[i] = cnfy[i] + qy[i];
  }
}

/* hydro_kn_getforceHG, hydro_kn_getforce.f90:73-107 */
static void k_getforceHG(const orc_state_t* s, int nel, double dt,
                         const int32_t* ielreg, const double* den,
                         const double* area, const double* cnu,
                         const double* cnv, double* cnfx, double* cnfy) {
  double w4 = 1.0 / dt;
  for (int iel = 1; iel <= nel; ++iel) {
    double w2 = CN(cnu, 1, iel) - CN(cnu, 2, iel) + CN(cnu, 3, iel) - CN(cnu, 4, iel);
    double w3 = CN(cnv, 1, iel) - CN(cnv, 2, iel) + CN(cnv, 3, iel) - CN(cnv, 4, iel);
    int ireg = EL(ielreg, iel);
    double w1 = -(s->kappareg[ireg - 1] * EL(den, iel) * EL(area, iel) * w4);
    w2 = w1 * w2;
    w3 = w1 * w3;
    CN(cnfx, 1, iel) = CN(cnfx, 1, iel) + w2;
    CN(cnfx, 2, iel) = CN(cnfx, 2, iel) - w2;
    CN(cnfx, 3, iel) = CN(cnfx, 3, iel) + w2;
    CN(cnfx, 4, iel) = CN(cnfx, 4, iel) - w2;
    CN(cnfy, 1, iel) = CN(cnfy, 1, iel) + w3;
    CN(cnfy, 2, iel) = CN(cnfy, 2, iel) - w3;
    CN(cnfy, 3, iel) = CN(cnfy, 3, iel) + w3;
    CN(cnfy, 4, iel) = CN(cnfy, 4, iel) - w3;
  }
}

/* hydro_kn_getforceSP, hydro_kn_getforce.f90:109-188 */
static void k_getforceSP(const orc_state_t* s, int nel, const int32_t* ielreg,
                         const double* den, const double* cs2,
                         const double* cnx, const double* cny,
                         const double* spmass, double* cnfx, double* cnfy) {
  for (int iel = 1; iel <= nel; ++iel) {
    int ireg = abs(EL(ielreg, iel));
    double w1 = s->pmeritreg[ireg - 1];
    double x3 = 0.25 * (CN(cnx, 1, iel) + CN(cnx, 2, iel) + CN(cnx, 3, iel) + CN(cnx, 4, iel));
    double y3 = 0.25 * (CN(cny, 1, iel) + CN(cny, 2, iel) + CN(cny, 3, iel) + CN(cny, 4, iel));
    double lfx[4][4], lfy[4][4]; /* lfx[j1-1][k-1] = lfx(j1,k) */
    for (int j1 = 1; j1 <= 4; ++j1) {
      double x1 = CN(cnx, j1, iel), y1 = CN(cny, j1, iel);
      int j2 = fmod_i(j1, 4) + 1;
      double x2 = 0.5 * (x1 + CN(cnx, j2, iel));
      double y2 = 0.5 * (y1 + CN(cny, j2, iel));
      j2 = fmod_i(j1 + 2, 4) + 1;
      double x4 = 0.5 * (x1 + CN(cnx, j2, iel));
      double y4 = 0.5 * (y1 + CN(cny, j2, iel));
      double w3 = 0.25 * (-x1 + x2 + x3 - x4);
      double w4 = 0.25 * (-x1 - x2 + x3 + x4);
      double w5 = 0.25 * (-y1 + y2 + y3 - y4);
      double w6 = 0.25 * (-y1 - y2 + y3 + y4);
      double w2 = 4.0 * (w3 * w6 - w4 * w5);
      w2 = CN(spmass, j1, iel) / w2;
      w2 = w2 - EL(den, iel);
      w2 = EL(cs2, iel) * w2;
      lfx[j1 - 1][0] = w2 * (w5 - w6);
      lfx[j1 - 1][1] = w2 * (w5 + w6);
      lfx[j1 - 1][2] = w2 * (-w5 + w6);
      lfx[j1 - 1][3] = w2 * (-w5 - w6);
      lfy[j1 - 1][0] = w2 * (-w3 + w4);
      lfy[j1 - 1][1] = w2 * (-w3 - w4);
      lfy[j1 - 1][2] = w2 * (w3 - w4);
      lfy[j1 - 1][3] = w2 * (w3 + w4);
    }
#define LF(a, i, k) a[(i)-1][(k)-1]
    double w2 = 0.5 * (LF(lfx, 1, 4) + LF(lfx, 4, 2));
    double w3 = 0.5 * (LF(lfx, 1, 2) + LF(lfx, 2, 4));
    double w4 = 0.5 * (LF(lfx, 2, 2) + LF(lfx, 3, 4));
    double w5 = 0.5 * (LF(lfx, 4, 4) + LF(lfx, 3, 2));
    double w6 = 0.25 * (LF(lfx, 1, 3) + LF(lfx, 2, 3) + LF(lfx, 3, 3) + LF(lfx, 4, 3));
    CN(cnfx, 1, iel) = CN(cnfx, 1, iel) + w1 * (LF(lfx, 1, 1) + w2 + w3 + w6);
    CN(cnfx, 2, iel) = CN(cnfx, 2, iel) + w1 * (LF(lfx, 2, 1) + w4 + w3 + w6);
    CN(cnfx, 3, iel) = CN(cnfx, 3, iel) + w1 * (LF(lfx, 3, 1) + w4 + w5 + w6);
    CN(cnfx, 4, iel) = CN(cnfx, 4, iel) + w1 * (LF(lfx, 4, 1) + w2 + w5 + w6);
    w2 = 0.5 * (LF(lfy, 1, 4) + LF(lfy, 4, 2));
    w3 = 0.5 * (LF(lfy, 1, 2) + LF(lfy, 2, 4));
    w4 = 0.5 * (LF(lfy, 2, 2) + LF(lfy, 3, 4));
    w5 = 0.5 * (LF(lfy, 4, 4) + LF(lfy, 3, 2));
    w6 = 0.25 * (LF(lfy, 1, 3) + LF(lfy, 2, 3) + LF(lfy, 3, 3) + LF(lfy, 4, 3));
    CN(cnfy, 1, iel) = CN(cnfy, 1, iel) + w1 * (LF(lfy, 1, 1) + w2 + w3 + w6);
    CN(cnfy, 2, iel) = CN(cnfy, 2, iel) + w1 * (LF(lfy, 2, 1) + w4 + w3 + w6);
    CN(cnfy, 3, iel) = CN(cnfy, 3, iel) + w1 * (LF(lfy, 3, 1) + w4 + w5 + w6);
    CN(cnfy, 4, iel) = CN(cnfy, 4, iel) + w1 * (LF(lfy, 4, 1) + w2 + w5 + w6);
#undef LF
  }
}

/* hydro_kn_viscinit / hydro_kn_spmassinit, hydro_kn_init.f90:29-77 */
static void k_spmassinit(int nel, const double* den, const double* cnx,
                         const double* cny, double* spmass) {
  for (int iel = 1; iel <= nel; ++iel) {
    double x3 = 0.25 * (CN(cnx, 1, iel) + CN(cnx, 2, iel) + CN(cnx, 3, iel) + CN(cnx, 4, iel));
    double y3 = 0.25 * (CN(cny, 1, iel) + CN(cny, 2, iel) + CN(cny, 3, iel) + CN(cny, 4, iel));
    for (int j1 = 1; j1 <= 4; ++j1) {
      double x1 = CN(cnx, j1, iel), y1 = CN(cny, j1, iel);
      int j2 = fmod_i(j1, 4) + 1;
      double x2 = 0.5 * (x1 + CN(cnx, j2, iel));
      double y2 = 0.5 * (y1 + CN(cny, j2, iel));
      j2 = fmod_i(j1 + 2, 4) + 1;
      double x4 = 0.5 * (x1 + CN(cnx, j2, iel));
      double y4 = 0.5 * (y1 + CN(cny, j2, iel));
      double w1 = 0.25 * (-x1 + x2 + x3 - x4);
      double w2 = 0.25 * (-x1 - x2 + x3 + x4);
      double w3 = 0.25 * (-y1 + y2 + y3 - y4);
      double w4 = 0.25 * (-y1 - y2 + y3 + y4);
      CN(spmass, j1, iel) = 4.0 * EL(den, iel) * (w1 * w4 - w2 * w3);
    }
  }
}

/* hydro_kn_scatteracceleration, hydro_kn_getacceleration.f90:30-74 */
static void k_scatteracceleration(int nel, int nnd, double zerocut,
                                  const int32_t* ielsort,
                                  const int32_t* ielnd, const double* den,
                                  const double* cnwt, const double* cnmass,
                                  double* ndarea, double* ndmass,
                                  const double* cnfx, const double* cnfy,
                                  double* ndudot, double* ndvdot) {
  for (int ind = 0; ind < nnd; ++ind) {
    ndmass[ind] = 0.0; ndarea[ind] = 0.0; ndudot[ind] = 0.0; ndvdot[ind] = 0.0;
  }
  for (int jj = 1; jj <= 4; ++jj) {
    for (int kk = 1; kk <= nel; ++kk) {
      int iel = EL(ielsort, kk);
      int ind = CN(ielnd, jj, iel);
      if (CN(cnmass, jj, iel) > zerocut) {
        EL(ndmass, ind) = EL(ndmass, ind) + CN(cnmass, jj, iel);
      } else {
        int ii = jj - 1;
        if (ii == 0) ii = 4;
        if (CN(cnmass, ii, iel) > zerocut)
          EL(ndmass, ind) = EL(ndmass, ind) + CN(cnmass, ii, iel);
        else
          EL(ndmass, ind) = EL(ndmass, ind) + EL(den, iel) * CN(cnwt, jj, iel);
      }
      EL(ndarea, ind) = EL(ndarea, ind) + CN(cnwt, jj, iel);
      EL(ndudot, ind) = EL(ndudot, ind) + CN(cnfx, jj, iel);
      EL(ndvdot, ind) = EL(ndvdot, ind) + CN(cnfy, jj, iel);
    }
  }
}

/* hydro_kn_getacceleration, hydro_kn_getacceleration.f90:76-100 */
static void k_getacceleration(int nnd, double dencut, double zerocut,
                              const double* ndarea, double* ndmass,
                              double* ndudot, double* ndvdot) {
  for (int ind = 0; ind < nnd; ++ind) {
    double w1 = dencut * ndarea[ind];
    if (ndmass[ind] > w1) {
      ndudot[ind] = ndudot[ind] / ndmass[ind];
      ndvdot[ind] = ndvdot[ind] / ndmass[ind];
    } else {
      ndudot[ind] = 0.0;
      ndvdot[ind] = 0.0;
      ndmass[ind] = fmax_f(zerocut, w1);
    }
  }
}

/* hydro_kn_applyacceleration, hydro_kn_getacceleration.f90:102-122 */
static void k_applyacceleration(int nnd, double dt, double* ndubar,
                                double* ndvbar, double* ndu, double* ndv) {
  double dt05 = 0.5 * dt;
  for (int ind = 0; ind < nnd; ++ind) {
    double w1 = ndu[ind], w2 = ndv[ind];
    ndu[ind] = w1 + dt * ndubar[ind];
    ndv[ind] = w2 + dt * ndvbar[ind];
    ndubar[ind] = w1 + dt05 * ndubar[ind];
    ndvbar[ind] = w2 + dt05 * ndvbar[ind];
  }
}

/* hydro_kn_getenergy, hydro_kn_getenergy.f90:29-52 */
static void k_getenergy(int nel, double dt, double zerocut, double* elenergy,
                        const double* elmass, const double* cnfx,
                        const double* cnfy, const double* cnu,
                        const double* cnv) {
  for (int iel = 1; iel <= nel; ++iel) {
    double w1 = CN(cnfx, 1, iel) * CN(cnu, 1, iel) + CN(cnfy, 1, iel) * CN(cnv, 1, iel) +
                CN(cnfx, 2, iel) * CN(cnu, 2, iel) + CN(cnfy, 2, iel) * CN(cnv, 2, iel) +
                CN(cnfx, 3, iel) * CN(cnu, 3, iel) + CN(cnfy, 3, iel) * CN(cnv, 3, iel) +
                CN(cnfx, 4, iel) * CN(cnu, 4, iel) + CN(cnfy, 4, iel) * CN(cnv, 4, iel);
    w1 = -(w1 / fmax_f(EL(elmass, iel), zerocut));
    EL(elenergy, iel) = EL(elenergy, iel) + w1 * dt;
  }
}

/* ===================================================================== */
/* ALE kernels                                                            */
/* ===================================================================== */

/* ale_kn_sumflux, ale_kn_advectors.f90:216-244 */
static void k_sumflux(int id1, int id2, int nel, const int32_t* ielel,
                      const int32_t* ielfc, const double* fcflux,
                      double* elflux) {
  for (int iel = 0; iel < nel; ++iel) elflux[iel] = 0.0;
  for (int i1 = id1; i1 <= id2; ++i1) {
    int i2 = i1 + 2;
    for (int iel = 1; iel <= nel; ++iel) {
      int iel1 = CN(ielel, i1, iel), iel2 = CN(ielel, i2, iel);
      int j1 = CN(ielfc, i1, iel), j2 = CN(ielfc, i2, iel);
      double w1 = CN(fcflux, j1, iel1);
      double w2 = CN(fcflux, j2, iel2);
      if (iel1 == iel) w1 = 0.0;
      if (iel2 == iel) w2 = 0.0;
      EL(elflux, iel) = EL(elflux, iel) - CN(fcflux, i1, iel) -
                        CN(fcflux, i2, iel) + w1 + w2;
    }
  }
}

/* ale_kn_fluxelvl, ale_kn_advectors.f90:30-81 */
static void k_fluxelvl(int id1, int id2, int nel, const int32_t* ielel,
                       const int32_t* ielfc, const double* cnbasis,
                       const double* fcdbasis, const double* elvar,
                       double* fcflux) {
  for (int64_t i = 0; i < 4 * (int64_t)nel; ++i) fcflux[i] = 0.0;
  for (int i1 = id1; i1 <= id2; ++i1) {
    int i2 = i1 + 2;
    for (int iel = 1; iel <= nel; ++iel) {
      int iel2 = CN(ielel, i2, iel);
      int j2 = CN(ielfc, i2, iel);
      int j1 = fmod_i(i2, 4) + 1;
      double r3 = CN(cnbasis, i2, iel) + CN(cnbasis, j1, iel);
      j1 = fmod_i(j2, 4) + 1;
      double w5 = r3 + CN(cnbasis, j2, iel2) + CN(cnbasis, j1, iel2);
      int iel1 = CN(ielel, i1, iel);
      j2 = CN(ielfc, i1, iel);
      j1 = i1 + 1;
      double r4 = CN(cnbasis, i1, iel) + CN(cnbasis, j1, iel);
      j1 = fmod_i(j2, 4) + 1;
      double w6 = r4 + CN(cnbasis, j2, iel1) + CN(cnbasis, j1, iel1);
      double rv = EL(elvar, iel);
      double r1 = CN(fcdbasis, i1, iel);
      double r2 = CN(fcdbasis, i2, iel);
      double w1 = rv - EL(elvar, iel2);
      double w2 = EL(elvar, iel1) - rv;
      double w3 = fabs(w1);
      double w4 = fabs(w2);
      double w7 = fsign(1.0, w2);
      double w8 = (w1 * w6 * w6 + w2 * w5 * w5) / (w5 * w6 * (w5 + w6));
      double grad = w7 * fmin_f(fmin_f(fabs(w8), w3 / w5), w4 / w6);
      if (w1 * w2 <= 0.0) grad = 0.0;
      r1 = r1 * (rv + grad * (r3 - 0.5 * r1));
      r2 = r2 * (rv - grad * (r4 - 0.5 * r2));
      CN(fcflux, i1, iel) = r1;
      CN(fcflux, i2, iel) = r2;
    }
  }
}

/* ale_kn_fluxndvl, ale_kn_advectors.f90:83-153 */
static void k_fluxndvl(int nel, const int32_t* ielel, const int32_t* ielfc,
                       const double* cnbasis, const double* cndbasis,
                       const double* cnvar, double* cnflux) {
  for (int64_t i = 0; i < 4 * (int64_t)nel; ++i) cnflux[i] = 0.0;
  for (int ilfcl = 1; ilfcl <= 2; ++ilfcl) {
    int ilfcr = ilfcl + 2;
    for (int icn = 1; icn <= 2; ++icn) {
      int ilndl = ilfcl + icn - 1;
      int ilndr = fmod_i(ilfcr - icn + 1, 4) + 1;
      for (int iel = 1; iel <= nel; ++iel) {
        double rd = 0.0;
        int iell = CN(ielel, ilfcl, iel), ielr = CN(ielel, ilfcr, iel);
        int ifcl = CN(ielfc, ilfcl, iel), ifcr = CN(ielfc, ilfcr, iel);
        int ii = ilfcl + 2 * (icn - 1);
        double cdb = CN(cndbasis, ii, iel);
        if (cdb > 0.0) {
          int ilnndl = fmod_i(ifcl + icn, 4) + 1;
          int ilnndr = fmod_i(ifcl - icn + 1, 4) + 1;
          double rv = CN(cnvar, ilndl, iel);
          rd = CN(cnbasis, ilndl, iel) - 0.5 * cdb;
          double w5 = CN(cnbasis, ilndl, iel) + CN(cnbasis, ilndr, iel);
          double w6 = CN(cnbasis, ilnndl, iell) + CN(cnbasis, ilnndr, iell);
          double w1 = rv - CN(cnvar, ilnndl, iell);
          double w2 = CN(cnvar, ilndr, iel) - rv;
          double w3 = fabs(w1), w4 = fabs(w2);
          double w7 = fsign(1.0, w2);
          double w8 = (w2 * w6 * w6 + w1 * w5 * w5) / (w5 * w6 * (w5 + w6));
          double grad = w7 * fmin_f(fmin_f(fabs(w8), w3 / w6), w4 / w5);
          if (w1 * w2 <= 0.0) grad = 0.0;
          rd = cdb * (rv + grad * rd);
        }
        if (cdb < 0.0) {
          int ilnndl = fmod_i(ifcr + icn - 2, 4) + 1;
          int ilnndr = fmodulo_i(ifcr - icn - 1, 4) + 1;
          double rv = CN(cnvar, ilndr, iel);
          rd = CN(cnbasis, ilndr, iel) + 0.5 * cdb;
          double w5 = CN(cnbasis, ilndl, iel) + CN(cnbasis, ilndr, iel);
          double w6 = CN(cnbasis, ilnndl, ielr) + CN(cnbasis, ilnndr, ielr);
          double w1 = rv - CN(cnvar, ilndl, iel);
          double w2 = CN(cnvar, ilnndr, ielr) - rv;
          double w3 = fabs(w1), w4 = fabs(w2);
          double w7 = fsign(1.0, w2);
          double w8 = (w1 * w6 * w6 + w2 * w5 * w5) / (w5 * w6 * (w5 + w6));
          double grad = -(w7 * fmin_f(fmin_f(fabs(w8), w3 / w5), w4 / w6));
          if (w1 * w2 <= 0.0) grad = 0.0;
          rd = cdb * (rv + grad * rd);
        }
        CN(cnflux, ilndl, iel) = CN(cnflux, ilndl, iel) - rd;
        CN(cnflux, ilndr, iel) = CN(cnflux, ilndr, iel) + rd;
      }
    }
  }
}

/* ale_kn_updateel, ale_kn_advectors.f90:155-180 */
static void k_updateel(int id1, int id2, int nel, const int32_t* ielel,
                       const int32_t* ielfc, const double* elbase0,
                       const double* elbase1, const double* cut,
                       const double* fcflux, double* elflux, double* elvar) {
  k_sumflux(id1, id2, nel, ielel, ielfc, fcflux, elflux);
  for (int iel = 0; iel < nel; ++iel)
    if (elbase1[iel] > cut[iel])
      elvar[iel] = (elvar[iel] * elbase0[iel] + elflux[iel]) / elbase1[iel];
}

/* ale_kn_updatend, ale_kn_advectors.f90:182-214 */
static void k_updatend(int nnd, int nel, const int32_t* ielnd,
                       const double* ndbase0, const double* ndbase1,
                       const double* cut, const uint8_t* zactive,
                       const double* cnflux, double* ndflux, double* ndvar) {
  for (int ind = 0; ind < nnd; ++ind) ndflux[ind] = 0.0;
  for (int iel = 1; iel <= nel; ++iel)
    for (int ii = 1; ii <= 4; ++ii) {
      int ind = CN(ielnd, ii, iel);
      EL(ndflux, ind) = EL(ndflux, ind) + CN(cnflux, ii, iel);
    }
  for (int ind = 0; ind < nnd; ++ind)
    if (zactive[ind] && (ndbase1[ind] > cut[ind]))
      ndvar[ind] = (ndvar[ind] * ndbase0[ind] + ndflux[ind]) / ndbase1[ind];
}

/* ale_kn_updatebasisel, ale_kn_advect.f90:31-61 */
static void k_updatebasisel(int nel, double zerocut, double dencut,
                            const double* totv, const double* totm,
                            double* cutv, double* cutm, double* elvpr,
                            double* elmpr, double* eldpr, double* elv,
                            double* elm, double* eld) {
  for (int iel = 0; iel < nel; ++iel) {
    elvpr[iel] = elv[iel];
    elmpr[iel] = elm[iel];
    eldpr[iel] = eld[iel];
    cutv[iel] = zerocut;
    cutm[iel] = elvpr[iel] * dencut;
    elv[iel] = elv[iel] + totv[iel];
    elm[iel] = elm[iel] + totm[iel];
    eld[iel] = elm[iel] / elv[iel];
  }
}

/* ale_kn_calcbasisnd, ale_kn_advect.f90:80-126 */
static void k_calcbasisnd(int nel, const int32_t* ielnd,
                          const int32_t* ielsort, const double* elv0,
                          const double* elv1, const double* cnm1,
                          double* ndv0, double* ndv1, double* ndm0,
                          double* cnm0) {
  for (int ii = 1; ii <= nel; ++ii) {
    int iel = EL(ielsort, ii);
    double w1 = 0.25 * EL(elv0, iel);
    double w2 = 0.25 * EL(elv1, iel);
    for (int j = 1; j <= 4; ++j) {
      double w3 = CN(cnm1, j, iel);
      CN(cnm0, j, iel) = w3;
      int ind = CN(ielnd, j, iel);
      EL(ndv0, ind) = EL(ndv0, ind) + w1;
      EL(ndv1, ind) = EL(ndv1, ind) + w2;
      EL(ndm0, ind) = EL(ndm0, ind) + w3;
    }
  }
}

/* ale_kn_fluxbasisnd, ale_kn_advect.f90:128-188 */
static void k_fluxbasisnd(int id1, int id2, int nel, const int32_t* ielel,
                          const int32_t* ielsd, const int32_t* ielsort,
                          const double* fcdv, const double* fcdm,
                          double* cndv, double* cndm, double* cnflux) {
  for (int64_t i = 0; i < 4 * (int64_t)nel; ++i) cnflux[i] = 0.0;
  for (int i1 = id1; i1 <= id2; ++i1) {
    int i2 = i1 + 2;
    for (int ii = 1; ii <= nel; ++ii) {
      int iel = EL(ielsort, ii);
      int ie1 = CN(ielel, i1, iel), ie2 = CN(ielel, i2, iel);
      int is1 = CN(ielsd, i1, iel), is2 = CN(ielsd, i2, iel);
      double w1 = CN(fcdv, is1, ie1), w2 = CN(fcdv, is2, ie2);
      double w3 = CN(fcdm, is1, ie1), w4 = CN(fcdm, is2, ie2);
      if (ie1 == iel) { w1 = 0.0; w3 = 0.0; }
      if (ie2 == iel) { w2 = 0.0; w4 = 0.0; }
      w1 = w1 - CN(fcdv, i1, iel);
      w2 = w2 - CN(fcdv, i2, iel);
      w1 = 0.25 * (w1 - w2);
      CN(cndv, i1, iel) = w1;
      CN(cndv, i2, iel) = w1;
      w1 = w3 - CN(fcdm, i1, iel);
      w2 = w4 - CN(fcdm, i2, iel);
      w3 = 0.25 * (w1 - w2);
      CN(cndm, i1, iel) = w3;
      CN(cndm, i2, iel) = w3;
      w3 = 0.25 * (w1 + w2);
      for (int j = 1; j <= 4; ++j) CN(cnflux, j, iel) = CN(cnflux, j, iel) + w3;
    }
  }
}

/* ale_kn_massbasisnd, ale_kn_advect.f90:190-219 */
static void k_massbasisnd(int nel, const int32_t* ielnd,
                          const int32_t* ielsort, const double* cnflux,
                          double* cnm1, double* ndm1) {
  for (int ii = 1; ii <= nel; ++ii) {
    int iel = EL(ielsort, ii);
    for (int j = 1; j <= 4; ++j)
      CN(cnm1, j, iel) = CN(cnm1, j, iel) + CN(cnflux, j, iel);
    for (int j = 1; j <= 4; ++j) {
      int ind = CN(ielnd, j, iel);
      EL(ndm1, ind) = EL(ndm1, ind) + CN(cnflux, j, iel);
    }
  }
}

/* ale_kn_activend, ale_kn_advect.f90:240-258 */
static void k_activend(int nnd, int ibc, const int32_t* indstatus,
                       const int32_t* indtype, uint8_t* zactive) {
  for (int ind = 0; ind < nnd; ++ind)
    zactive[ind] = (indstatus[ind] > 0 && indtype[ind] != ibc &&
                    indtype[ind] != -3) ? 1 : 0;
}

/* ===================================================================== */
/* drivers                                                                */
/* ===================================================================== */

/* geometry_dr_getgeometry, geometry_dr_volume.f90:40-96 (single material) */
/* ---- mixed cells: component gathers and averages ------------------------
 * The flattened lists (mix_dr_flatten) keep each mixed cell's components
 * contiguous from imxfcp(imx), imxncp(imx) of them; every loop below walks
 * them in that order as the reference does. */

/* utils_kn_mxgather1/4, utils_kn_gather.f90:53-106: element value to each of
 * the cell's components */
static void k_mxgather(int nmx, const int32_t* imxel, const int32_t* imxfcp,
                       const int32_t* imxncp, int narr, const double** el,
                       double** mx) {
  for (int imx = 1; imx <= nmx; ++imx) {
    int iel = EL(imxel, imx), icp = EL(imxfcp, imx);
    for (int ii = 1; ii <= EL(imxncp, imx); ++ii) {
      for (int k = 0; k < narr; ++k) EL(mx[k], icp) = EL(el[k], iel);
      icp = icp + 1;
    }
  }
}

/* utils_kn_mxgathercnav, utils_kn_gather.f90:130-152:
 * mx(:,icp) = frac(icp)*el(:,iel) */
static void k_mxgathercnav(int nmx, const int32_t* imxel,
                           const int32_t* imxfcp, const int32_t* imxncp,
                           const double* el, const double* frac, double* mx) {
  for (int imx = 1; imx <= nmx; ++imx) {
    int iel = EL(imxel, imx), icp = EL(imxfcp, imx);
    double w1[4];
    for (int j = 1; j <= 4; ++j) w1[j - 1] = CN(el, j, iel);
    for (int ii = 1; ii <= EL(imxncp, imx); ++ii) {
      for (int j = 1; j <= 4; ++j) CN(mx, j, icp) = EL(frac, icp) * w1[j - 1];
      icp = icp + 1;
    }
  }
}

/* utils_kn_average12, utils_kn_average.f90:84-105:
 * el(iel) = sum frac*(mx1+mx2) over the cell's components */
static void k_average12(int nmx, const int32_t* imxel, const int32_t* imxfcp,
                        const int32_t* imxncp, const double* frac,
                        const double* mx1, const double* mx2, double* el) {
  for (int imx = 1; imx <= nmx; ++imx) {
    double w1 = 0.0;
    int icp = EL(imxfcp, imx);
    for (int ii = 1; ii <= EL(imxncp, imx); ++ii) {
      w1 = w1 + EL(frac, icp) * (EL(mx1, icp) + EL(mx2, icp));
      icp = icp + 1;
    }
    EL(el, EL(imxel, imx)) = w1;
  }
}

/* utils_kn_average22, utils_kn_average.f90:53-82: corner arrays */
static void k_average22(int nmx, const int32_t* imxel, const int32_t* imxfcp,
                        const int32_t* imxncp, const double* frac,
                        const double* mx1, const double* mx2, double* el1,
                        double* el2) {
  for (int imx = 1; imx <= nmx; ++imx) {
    double w1[4] = {0.0, 0.0, 0.0, 0.0}, w2[4] = {0.0, 0.0, 0.0, 0.0};
    int icp = EL(imxfcp, imx);
    for (int ii = 1; ii <= EL(imxncp, imx); ++ii) {
      for (int j = 1; j <= 4; ++j) {
        w1[j - 1] = w1[j - 1] + EL(frac, icp) * CN(mx1, j, icp);
        w2[j - 1] = w2[j - 1] + EL(frac, icp) * CN(mx2, j, icp);
      }
      icp = icp + 1;
    }
    int iel = EL(imxel, imx);
    for (int j = 1; j <= 4; ++j) {
      CN(el1, j, iel) = w1[j - 1];
      CN(el2, j, iel) = w2[j - 1];
    }
  }
}

#define MXLISTS I(BL_IMXEL), I(BL_IMXFCP), I(BL_IMXNCP)

int orc_geometry_getgeometry(orc_state_t* s, bl_error_t* err) {
  int nel = s->sz.nel;
  k_cngather(nel, I(BL_IELND), R(BL_NDX), R(BL_CNX));
  k_cngather(nel, I(BL_IELND), R(BL_NDY), R(BL_CNY));
  k_getiso(nel, R(BL_CNX), R(BL_CNY), R(BL_A1), R(BL_A2), R(BL_A3), R(BL_B1),
           R(BL_B2), R(BL_B3), R(BL_CNWT));
  k_getvolume(nel, R(BL_A1), R(BL_A3), R(BL_B1), R(BL_B3), R(BL_ELVOLUME));
  int ncp = s->sz.ncp;
  if (ncp > 0) { /* geometry_dr_volume.f90:66-79 */
    const double* el[4] = {R(BL_A1), R(BL_A3), R(BL_B1), R(BL_B3)};
    double* mx[4] = {R(BL_CPA1), R(BL_CPA3), R(BL_CPB1), R(BL_CPB3)};
    k_mxgather(s->sz.nmx, MXLISTS, 4, el, mx);
    k_getvolume(ncp, R(BL_CPA1), R(BL_CPA3), R(BL_CPB1), R(BL_CPB3),
                R(BL_CPVOLUME));
    for (int icp = 1; icp <= ncp; ++icp) /* utils_kn_multiply */
      EL(R(BL_CPVOLUME), icp) = EL(R(BL_CPVOLUME), icp) * EL(R(BL_FRVOLUME), icp);
  }
  int ierr = k_checkvolume(nel, 0.0, R(BL_ELVOLUME));
  if (ierr != BL_SUCCESS) {
    set_err(err, ierr, BL_HALT_SINGLE,
            "ERROR: negative volume in geometry_dr_getgeometry");
    return BL_FAILURE;
  }
  set_err(err, BL_SUCCESS, 0, "");
  return BL_SUCCESS;
}

/* density_dr_getdensity, density_dr_getdensity.f90:32-46 */
int orc_density_getdensity(orc_state_t* s) {
  k_divide(s->sz.nel, R(BL_ELMASS), R(BL_ELVOLUME), R(BL_ELDENSITY));
  if (s->sz.ncp > 0)
    k_divide(s->sz.ncp, R(BL_CPMASS), R(BL_CPVOLUME), R(BL_CPDENSITY));
  return BL_SUCCESS;
}

/* init_kn_connectivity (init_kn.f90:82-91): ielel = element across face f
 * (face f joins corners f and f+1), ielfc = the face of that neighbour which
 * points back (the LAST such face, as utils_kn_getsconn's istore scan leaves
 * it, utils_kn_connectivity.f90:113-134); boundary faces get ielel = self and
 * ielfc = the opposite face (utils_kn_corrconn, :136-168).
 * init_kn_nodetype (init_kn.f90:93-113) and ielsort = identity
 * (init_dr_utils.f90:64-69).  Brute force over the elements of each node. */
int orc_init_mesh(orc_state_t* s) {
  int nel = s->sz.nel, nnd = s->sz.nnd;
  int32_t *ielnd = I(BL_IELND), *ielel = I(BL_IELEL), *ielfc = I(BL_IELFC);
  int32_t* indtype = I(BL_INDTYPE);
  /* node -> elements lists */
  int* cnt = (int*)calloc(nnd + 1, sizeof(int));
  for (int64_t k = 0; k < 4 * (int64_t)nel; ++k) cnt[ielnd[k]] += 1;
  int* start = (int*)calloc(nnd + 2, sizeof(int));
  for (int n = 1; n <= nnd; ++n) start[n + 1] = start[n] + cnt[n];
  int* fillp = (int*)calloc(nnd + 2, sizeof(int));
  memcpy(fillp, start, sizeof(int) * (nnd + 2));
  int* lst = (int*)malloc(sizeof(int) * 4 * (size_t)nel);
  for (int e = 1; e <= nel; ++e)
    for (int j = 1; j <= 4; ++j) lst[fillp[CN(ielnd, j, e)]++] = e;
  for (int e = 1; e <= nel; ++e) {
    for (int f = 1; f <= 4; ++f) {
      int a = CN(ielnd, f, e), b = CN(ielnd, (f % 4) + 1, e);
      int nb = 0;
      for (int k = start[a]; k < start[a + 1]; ++k) {
        int o = lst[k];
        if (o == e) continue;
        for (int g = 1; g <= 4; ++g) {
          int c = CN(ielnd, g, o), d = CN(ielnd, (g % 4) + 1, o);
          if ((c == a && d == b) || (c == b && d == a)) nb = o;
        }
      }
      CN(ielel, f, e) = nb;
    }
  }
  for (int e = 1; e <= nel; ++e)
    for (int f = 1; f <= 4; ++f) {
      int nb = CN(ielel, f, e);
      int sd = 0;
      if (nb > 0)
        for (int g = 1; g <= 4; ++g)
          if (CN(ielel, g, nb) == e) sd = g;
      CN(ielfc, f, e) = sd;
    }
  for (int e = 1; e <= nel; ++e)
    for (int f = 1; f <= 2; ++f) {
      int f2 = f + 2;
      if (CN(ielel, f, e) == 0) { CN(ielel, f, e) = e; CN(ielfc, f, e) = f2; }
      if (CN(ielel, f2, e) == 0) { CN(ielel, f2, e) = e; CN(ielfc, f2, e) = f; }
    }
  /* init_kn_nodetype */
  for (int e = 1; e <= nel; ++e) {
    int nodes[4], c = 0;
    for (int j = 0; j < 4; ++j) {
      nodes[j] = CN(ielnd, j + 1, e);
      if (EL(indtype, nodes[j]) < 0) ++c;
    }
    if (c != 3) continue;
    int ii;
    for (ii = 0; ii < 4; ++ii) if (EL(indtype, nodes[ii]) > 0) break;
    ii = (ii + 2) % 4;
    int jj = nodes[ii];
    if (jj <= nnd) {
      int j1 = nodes[(ii + 1) % 4], j2 = nodes[(ii + 3) % 4];
      if ((EL(indtype, j1) == -2 && EL(indtype, j2) == -1) ||
          (EL(indtype, j2) == -2 && EL(indtype, j1) == -1))
        EL(indtype, jj) = -3;
    }
  }
  for (int e = 1; e <= nel; ++e) {
    EL(I(BL_IELSORT1), e) = e;
    EL(I(BL_IELSORT2), e) = e;
  }
  free(cnt); free(start); free(fillp); free(lst);
  return BL_SUCCESS;
}

/* init_dr_state, init_dr_utils.f90:93-117 */
int orc_init_state(orc_state_t* s) {
  int nel = s->sz.nel;
  double *den = R(BL_ELDENSITY), *vol = R(BL_ELVOLUME), *cnwt = R(BL_CNWT);
  double *m = R(BL_ELMASS), *cm = R(BL_CNMASS);
  for (int iel = 1; iel <= nel; ++iel) { /* init_kn_elmass, init_kn.f90:32-48 */
    EL(m, iel) = EL(den, iel) * EL(vol, iel);
    for (int j = 1; j <= 4; ++j) CN(cm, j, iel) = EL(den, iel) * CN(cnwt, j, iel);
  }
  int ncp = s->sz.ncp;
  if (ncp > 0) { /* init_dr_mass, init_dr_utils.f90:121-130 */
    for (int icp = 1; icp <= ncp; ++icp) /* init_kn_mxmass */
      EL(R(BL_CPMASS), icp) = EL(R(BL_CPDENSITY), icp) * EL(R(BL_CPVOLUME), icp);
    const double* el[1] = {m};
    double* mx[1] = {R(BL_RCPSCRATCH11)};
    k_mxgather(s->sz.nmx, MXLISTS, 1, el, mx);
    k_divide(ncp, R(BL_CPMASS), R(BL_RCPSCRATCH11), R(BL_FRMASS));
  }
  return orc_hydro_geteos(s);
}

/* hydro_dr_init, hydro_dr_init.f90:37-57 */
int orc_hydro_init(orc_state_t* s) {
  int nel = s->sz.nel2;
  for (int iel = 1; iel <= nel; ++iel) {
    EL(R(BL_ELVISC), iel) = 0.0;
    for (int j = 1; j <= 4; ++j) {
      CN(R(BL_CNVISCX), j, iel) = 0.0;
      CN(R(BL_CNVISCY), j, iel) = 0.0;
    }
  }
  for (int icp = 1; icp <= s->sz.ncp; ++icp) { /* hydro_dr_init.f90:48-51 */
    EL(R(BL_CPVISC), icp) = 0.0;
    for (int j = 1; j <= 4; ++j) {
      CN(R(BL_CPVISCX), j, icp) = 0.0;
      CN(R(BL_CPVISCY), j, icp) = 0.0;
    }
  }
  if (s->cfg.zsp)
    k_spmassinit(s->sz.nel, R(BL_ELDENSITY), R(BL_CNX), R(BL_CNY), R(BL_SPMASS));
  return BL_SUCCESS;
}

/* hydro_dr_getdt, hydro_dr_getdt.f90:39-79 with hydro_dr_getviscosity
 * (hydro_dr_getviscosity.f90:41-91) and hydro_dr_getcs2 */
int orc_hydro_getdt(orc_state_t* s, bl_dt_t* cfl, bl_dt_t* div,
                    bl_error_t* err) {
  int nel = s->sz.nel;
  const bl_config_t* c = &s->cfg;
  /* dtstep aliases, dataAPI_id.f90:78-84 */
  double *cnu = R(BL_RSCRATCH21), *cnv = R(BL_RSCRATCH22);
  double *du = R(BL_RSCRATCH23), *dv = R(BL_RSCRATCH24);
  double *dx = R(BL_RSCRATCH25), *dy = R(BL_RSCRATCH26);
  double *store = R(BL_RSCRATCH27);
  double *scratch = R(BL_RSCRATCH11), *ellength = R(BL_RSCRATCH12);
  k_cngather(nel, I(BL_IELND), R(BL_NDU), cnu);
  k_cngather(nel, I(BL_IELND), R(BL_NDV), cnv);
  k_initeviscosity(nel, R(BL_CNX), R(BL_CNY), cnu, cnv, R(BL_ELVISC), dx, dy,
                   du, dv, R(BL_CNVISCX), R(BL_CNVISCY));
  k_limiteviscosity(nel, c->zerocut, c->cvisc1, c->cvisc2, I(BL_INDTYPE),
                    I(BL_IELEL), I(BL_IELND), I(BL_IELFC), R(BL_ELDENSITY),
                    R(BL_ELCS2), du, dv, dx, dy, store, R(BL_CNVISCX),
                    R(BL_CNVISCY), R(BL_ELVISC));
  k_geteviscosity(nel, c->zerocut, R(BL_CNX), R(BL_CNY), cnu, cnv,
                  R(BL_CNVISCX), R(BL_CNVISCY));
  k_sum(nel, R(BL_ELVISC), R(BL_ELCS2));
  if (s->sz.ncp > 0) /* hydro_dr_getcs2.f90:40-42 */
    k_average12(s->sz.nmx, MXLISTS, R(BL_FRMASS), R(BL_CPCS2), R(BL_CPVISC),
                R(BL_ELCS2));
  double rdt = 0.0;
  int idt = 0;
  int ierr = k_getdtcfl(s, nel, c->zcut, c->cfl_sf, I(BL_IELREG), R(BL_ELCS2),
                        R(BL_CNX), R(BL_CNY), scratch, ellength, &rdt, &idt);
  if (ierr != BL_SUCCESS) {
    set_err(err, ierr, BL_HALT_SINGLE,
            "ERROR: negative CFL condition in hydro_dr_getdt");
    return BL_FAILURE;
  }
  cfl->rdt = rdt; cfl->idt = idt; cfl->sdt = BL_DT_CFL;
  k_getdtdiv(nel, c->div_sf, R(BL_A1), R(BL_A3), R(BL_B1), R(BL_B3),
             R(BL_ELVOLUME), cnu, cnv, &rdt, &idt);
  div->rdt = rdt; div->idt = idt; div->sdt = BL_DT_DIV;
  set_err(err, BL_SUCCESS, 0, "");
  return BL_SUCCESS;
}

/* hydro_dr_set.f90:31-51 */
int orc_hydro_setpredictor(orc_state_t* s) {
  k_copy(s->sz.nel, R(BL_ELENERGY), R(BL_RSCRATCH11));
  return BL_SUCCESS;
}
int orc_hydro_setcorrector(orc_state_t* s) {
  k_copy(s->sz.nel, R(BL_RSCRATCH11), R(BL_ELENERGY));
  return BL_SUCCESS;
}

/* hydro_dr_getforce, hydro_dr_getforce.f90:40-91; component forces cpfx/cpfy
 * are rcpscratch21/22 (dataAPI_id.f90:94-96) */
int orc_hydro_getforce(orc_state_t* s, const bl_timestep_t* ts) {
  int nel = s->sz.nel, ncp = s->sz.ncp;
  double *cnfx = R(BL_RSCRATCH23), *cnfy = R(BL_RSCRATCH24);
  k_getforceP(nel, R(BL_ELPRESSURE), R(BL_A1), R(BL_A3), R(BL_B1), R(BL_B3),
              cnfx, cnfy);
  if (ncp > 0)
    k_getforceP(ncp, R(BL_CPPRESSURE), R(BL_CPA1), R(BL_CPA3), R(BL_CPB1),
                R(BL_CPB3), R(BL_RCPSCRATCH21), R(BL_RCPSCRATCH22));
  k_getforceQ(nel, R(BL_CNVISCX), R(BL_CNVISCY), cnfx, cnfy);
  if (ncp > 0)
    k_getforceQ(ncp, R(BL_CPVISCX), R(BL_CPVISCY), R(BL_RCPSCRATCH21),
                R(BL_RCPSCRATCH22));
  if (s->cfg.zsp)
    k_getforceSP(s, nel, I(BL_IELREG), R(BL_ELDENSITY), R(BL_ELCS2), R(BL_CNX),
                 R(BL_CNY), R(BL_SPMASS), cnfx, cnfy);
  if (ts->zcorrector && s->cfg.zhg)
    k_getforceHG(s, nel, ts->dt, I(BL_IELREG), R(BL_ELDENSITY),
                 R(BL_ELVOLUME), R(BL_RSCRATCH21), R(BL_RSCRATCH22), cnfx,
                 cnfy);
  return BL_SUCCESS;
}

/* hydro_dr_getstate, hydro_dr_getstate.f90:33-51 */
int orc_hydro_getstate(orc_state_t* s, const bl_timestep_t* ts,
                       bl_error_t* err) {
  k_getvertex(s->sz.nnd, 0.5 * ts->dt, R(BL_NDX), R(BL_NDY), R(BL_NDU),
              R(BL_NDV));
  if (orc_geometry_getgeometry(s, err) != BL_SUCCESS) return BL_FAILURE;
  return orc_density_getdensity(s);
}

/* hydro_dr_getenergy, hydro_dr_getenergy.f90:35-64 */
int orc_hydro_getenergy(orc_state_t* s, const bl_timestep_t* ts) {
  k_getenergy(s->sz.nel, ts->dts, s->cfg.zerocut, R(BL_ELENERGY), R(BL_ELMASS),
              R(BL_RSCRATCH23), R(BL_RSCRATCH24), R(BL_RSCRATCH21),
              R(BL_RSCRATCH22));
  if (s->sz.ncp > 0) {
    /* hydro_dr_getenergy.f90:47-62: the component corner velocities cpu/cpv
     * are rscratch23/24, the same memory as the element corner forces
     * (dataAPI_id.f90:94-96), written after the element update has read them */
    k_mxgathercnav(s->sz.nmx, MXLISTS, R(BL_RSCRATCH21), R(BL_FRVOLUME),
                   R(BL_RSCRATCH23));
    k_mxgathercnav(s->sz.nmx, MXLISTS, R(BL_RSCRATCH22), R(BL_FRVOLUME),
                   R(BL_RSCRATCH24));
    k_getenergy(s->sz.ncp, ts->dts, s->cfg.zerocut, R(BL_CPENERGY),
                R(BL_CPMASS), R(BL_RCPSCRATCH21), R(BL_RCPSCRATCH22),
                R(BL_RSCRATCH23), R(BL_RSCRATCH24));
  }
  return BL_SUCCESS;
}

/* hydro_dr_geteos -> eos_dr_geteos, eos_dr_geteos.f90:35-66 */
int orc_hydro_geteos(orc_state_t* s) {
  k_geteos(s, s->sz.nel, I(BL_IELMAT), R(BL_ELDENSITY), R(BL_ELENERGY),
           R(BL_ELPRESSURE), R(BL_ELCS2));
  if (s->sz.ncp > 0)
    k_geteos(s, s->sz.ncp, I(BL_ICPMAT), R(BL_CPDENSITY), R(BL_CPENERGY),
             R(BL_CPPRESSURE), R(BL_CPCS2));
  return BL_SUCCESS;
}

/* hydro_dr_getacceleration.f90:44-112 */
int orc_hydro_initacceleration(orc_state_t* s) {
  /* mixed cells: the element corner forces become the volume-fraction
   * average of the component forces (hydro_dr_getacceleration.f90:69-72);
   * nproc=1: no exchange */
  if (s->sz.ncp > 0)
    k_average22(s->sz.nmx, MXLISTS, R(BL_FRVOLUME), R(BL_RCPSCRATCH21),
                R(BL_RCPSCRATCH22), R(BL_RSCRATCH23), R(BL_RSCRATCH24));
  return BL_SUCCESS;
}
int orc_hydro_scatteracceleration(orc_state_t* s) {
  k_scatteracceleration(s->sz.nel1, s->sz.nnd1, s->cfg.zerocut,
                        I(BL_IELSORT1), I(BL_IELND), R(BL_ELDENSITY),
                        R(BL_CNWT), R(BL_CNMASS), R(BL_RSCRATCH12),
                        R(BL_RSCRATCH13), R(BL_RSCRATCH23), R(BL_RSCRATCH24),
                        R(BL_RSCRATCH14), R(BL_RSCRATCH15));
  return BL_SUCCESS;
}
int orc_hydro_getacceleration(orc_state_t* s) {
  k_getacceleration(s->sz.nnd, s->cfg.dencut, s->cfg.zerocut, R(BL_RSCRATCH12),
                    R(BL_RSCRATCH13), R(BL_RSCRATCH14), R(BL_RSCRATCH15));
  k_bc(s->sz.nnd, s->cfg.accut, I(BL_INDTYPE), R(BL_RSCRATCH14),
       R(BL_RSCRATCH15));
  return BL_SUCCESS;
}
int orc_hydro_applyacceleration(orc_state_t* s, const bl_timestep_t* ts) {
  k_applyacceleration(s->sz.nnd, ts->dt, R(BL_RSCRATCH14), R(BL_RSCRATCH15),
                      R(BL_NDU), R(BL_NDV));
  k_cngather(s->sz.nel, I(BL_IELND), R(BL_RSCRATCH14), R(BL_RSCRATCH21));
  k_cngather(s->sz.nel, I(BL_IELND), R(BL_RSCRATCH15), R(BL_RSCRATCH22));
  return BL_SUCCESS;
}

/* ale_active, ale_dr_utils.f90:35-51 */
int orc_ale_active(orc_state_t* s, const bl_timestep_t* ts) {
  if (!s->cfg.zexist) return 0;
  return (ts->time >= s->cfg.mintime) && (ts->time <= s->cfg.maxtime);
}

/* ---- node relaxation ------------------------------------------------ */

/* The 8-node stencil of each node, in the order of the weights r1..r8 of
 * setup_mesh_region (setup_mesh.F90:848-857):
 *   (lp,km) (lp,kk) (lp,kp) (ll,km) (ll,kp) (lm,km) (lm,kk) (lm,kp).
 * Frame: e0 is the lowest-numbered element with n at corner 1, else the
 * lowest-numbered element around n; with n at corner j of e0, +k points to
 * corner j+1 and +l to corner j+3 (counter-clockwise corners).  Walking
 * counter-clockwise around n from +k: the next element e' after e has its
 * corner after n equal to e's corner before n. */
static void build_ring(orc_state_t* s) {
  int nel = s->sz.nel, nnd = s->sz.nnd;
  const int32_t *ielnd = I(BL_IELND), *indtype = I(BL_INDTYPE);
  free(s->ring);
  s->ring = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)(nnd > 0 ? nnd : 1));
  for (int64_t k = 0; k < 8 * (int64_t)nnd; ++k) s->ring[k] = -1;
  int* cnt = (int*)calloc(nnd + 1, sizeof(int));
  for (int64_t k = 0; k < 4 * (int64_t)nel; ++k) cnt[ielnd[k] - 1] += 1;
  int* start = (int*)calloc(nnd + 1, sizeof(int));
  for (int n = 0; n < nnd; ++n) start[n + 1] = start[n] + cnt[n];
  int* fill = (int*)malloc(sizeof(int) * (nnd + 1));
  memcpy(fill, start, sizeof(int) * (nnd + 1));
  int* lst = (int*)malloc(sizeof(int) * 4 * (size_t)(nel > 0 ? nel : 1));
  for (int e = 0; e < nel; ++e) /* ascending element order per node */
    for (int j = 0; j < 4; ++j) {
      int n = ielnd[4 * (int64_t)e + j] - 1;
      lst[fill[n]++] = 4 * e + j;
    }
  for (int n = 0; n < nnd; ++n) {
    if (indtype[n] <= 0 || start[n + 1] - start[n] != 4) continue;
    int c0 = lst[start[n]];
    for (int k = start[n]; k < start[n + 1]; ++k)
      if ((lst[k] & 3) == 0) { c0 = lst[k]; break; }
#define NODE_AT(c, d) (ielnd[4 * (int64_t)((c) >> 2) + (((c) + (d)) & 3)] - 1)
    int P[8];
    int used[4] = {c0, -1, -1, -1};
    P[0] = NODE_AT(c0, 1);
    P[1] = NODE_AT(c0, 2);
    P[2] = NODE_AT(c0, 3);
    int ok = 1, cur = c0;
    for (int step = 1; step < 4 && ok; ++step) {
      int prev = NODE_AT(cur, 3), nxt = -1;
      for (int k = start[n]; k < start[n + 1]; ++k) {
        int c = lst[k];
        if (c == used[0] || c == used[1] || c == used[2]) continue;
        if (NODE_AT(c, 1) == prev) { nxt = c; break; }
      }
      if (nxt < 0) { ok = 0; break; }
      used[step] = nxt;
      cur = nxt;
      P[2 * step + 1] = NODE_AT(cur, 2);
      if (step < 3) P[2 * step + 2] = NODE_AT(cur, 3);
    }
    if (!ok || NODE_AT(cur, 3) != P[0]) continue; /* ring not closed */
#undef NODE_AT
    /* P: counter-clockwise from +k: (ll,kp) (lp,kp) (lp,kk) (lp,km)
     * (ll,km) (lm,km) (lm,kk) (lm,kp) */
    int32_t* r = s->ring + 8 * (int64_t)n;
    r[0] = P[3]; r[1] = P[2]; r[2] = P[1]; r[3] = P[4];
    r[4] = P[0]; r[5] = P[5]; r[6] = P[6]; r[7] = P[7];
  }
  free(cnt); free(start); free(fill); free(lst);
}

/* one node of one sweep (setup_mesh.F90:797-863): x (ss) and y (rr) read
 * from (xi, yi), the result written to (xo, yo) */
static void relax_node(int type, double om, double tol, double zerocut, int plant,
                       const int32_t* nb, const double* xi, const double* yi,
                       int n, double* xo, double* yo) {
  double xn = xi[n], yn = yi[n];
  xo[n] = xn;
  yo[n] = yn;
  if (nb[0] < 0) return;
  double r1, r2, r3, r4, r5, r6, r7, r8;
  if (type == BL_RELAX_EQUI) {
    /* (ll,kp) is nb[4], (ll,km) nb[3], (lp,kk) nb[1], (lm,kk) nb[6] */
    double r_psi = 0.5 * (yi[nb[4]] - yi[nb[3]]);
    double s_psi = 0.5 * (xi[nb[4]] - xi[nb[3]]);
    double r_phi = 0.5 * (yi[nb[1]] - yi[nb[6]]);
    double s_phi = 0.5 * (xi[nb[1]] - xi[nb[6]]);
    double zeta = s_psi * r_phi - s_phi * r_psi;
    if (fabs(zeta) < zerocut) return;
    double alph = s_psi * s_psi + r_psi * r_psi;
    double beta = s_phi * s_psi + r_phi * r_psi;
    double gamm = s_phi * s_phi + r_phi * r_phi;
    if (plant == 1) { double t = alph; alph = gamm; gamm = t; }
    if (plant == 2) beta = -beta;
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
    /* LIN1: r_wgt(2) = r_wgt(7) = 1, LIN2: r_wgt(4) = r_wgt(5) = 1, each
     * divided by their sum (setup_mesh.F90:626-660) */
    double w = 1.0 / 2.0;
    r1 = 0.0; r2 = 0.0; r3 = 0.0; r4 = 0.0; r5 = 0.0; r6 = 0.0; r7 = 0.0; r8 = 0.0;
    if (type == BL_RELAX_LIN1) { r2 = w; r7 = w; } else { r4 = w; r5 = w; }
  }
  /* s_i = r_i for every type here */
  double dr = r1 * yi[nb[0]] + r2 * yi[nb[1]] + r3 * yi[nb[2]] + r4 * yi[nb[3]] +
              r5 * yi[nb[4]] + r6 * yi[nb[5]] + r7 * yi[nb[6]] + r8 * yi[nb[7]] - yn;
  double ds = r1 * xi[nb[0]] + r2 * xi[nb[1]] + r3 * xi[nb[2]] + r4 * xi[nb[3]] +
              r5 * xi[nb[4]] + r6 * xi[nb[5]] + r7 * xi[nb[6]] + r8 * xi[nb[7]] - xn;
  if ((fabs(dr) > tol) || (fabs(ds) > tol)) {
    yo[n] = yn + om * dr;
    xo[n] = xn + om * ds;
  }
}

/* nsweep Jacobi sweeps from (x, y) into (xr, yr) */
static void relax_sweeps(orc_state_t* s, int type, int nsweep, double om, double tol,
                         const double* x, const double* y, double* xr, double* yr) {
  int nnd = s->sz.nnd;
  if (!s->ring) build_ring(s);
  double* tx = (double*)malloc(sizeof(double) * (nnd > 0 ? nnd : 1));
  double* ty = (double*)malloc(sizeof(double) * (nnd > 0 ? nnd : 1));
  memcpy(xr, x, sizeof(double) * nnd);
  memcpy(yr, y, sizeof(double) * nnd);
  for (int k = 0; k < nsweep; ++k) {
    memcpy(tx, xr, sizeof(double) * nnd);
    memcpy(ty, yr, sizeof(double) * nnd);
    for (int n = 0; n < nnd; ++n)
      relax_node(type, om, tol, s->cfg.zerocut, s->relax_plant, s->ring + 8 * (int64_t)n, tx,
                 ty, n, xr, yr);
  }
  free(tx); free(ty);
}

static int relax_args_ok(int32_t type, int32_t nsweep, double om, double tol) {
  return type >= BL_RELAX_OFF && type <= BL_RELAX_EQUI && nsweep >= 1 && om > 0.0 &&
         tol >= 0.0;
}

int orc_set_relaxation(orc_state_t* s, int32_t type, int32_t nsweep, double om,
                       double tol) {
  if (!relax_args_ok(type, nsweep, om, tol)) return BL_FAILURE;
  s->relax_type = type;
  s->relax_nsweep = nsweep;
  s->relax_om = om;
  s->relax_tol = tol;
  build_ring(s);
  return BL_SUCCESS;
}

int orc_relax_mesh(orc_state_t* s, int32_t type, int32_t nsweep, double om, double tol) {
  if (!relax_args_ok(type, nsweep, om, tol) || type == BL_RELAX_OFF) return BL_FAILURE;
  int nnd = s->sz.nnd;
  build_ring(s);
  double* xr = (double*)malloc(sizeof(double) * (nnd > 0 ? nnd : 1));
  double* yr = (double*)malloc(sizeof(double) * (nnd > 0 ? nnd : 1));
  relax_sweeps(s, type, nsweep, om, tol, R(BL_NDX), R(BL_NDY), xr, yr);
  memcpy(R(BL_NDX), xr, sizeof(double) * nnd);
  memcpy(R(BL_NDY), yr, sizeof(double) * nnd);
  free(xr); free(yr);
  return BL_SUCCESS;
}

/* tests only: plant a mistake in the EQUI stencil (1: alpha and gamma
 * swapped, 2: beta of the wrong sign; 0: none), to show that the tests of
 * the stencil detect it */
void orc_relax_plant(orc_state_t* s, int32_t which) { s->relax_plant = which; }

/* the stencil table (8, nnd), for tests */
const int32_t* orc_relax_ring(orc_state_t* s) {
  if (!s->ring) build_ring(s);
  return s->ring;
}

/* ale_dr_getmeshstatus, ale_dr_getmeshstatus.f90:32-49 (relaxation marks
 * every node moving, as zeul does) */
int orc_ale_getmeshstatus(orc_state_t* s) {
  if (s->cfg.zeul || s->relax_type != BL_RELAX_OFF) {
    int32_t* st = I(BL_ISCRATCH11);
    for (int ind = 0; ind < s->sz.nnd; ++ind) st[ind] = 2;
  }
  return BL_SUCCESS;
}

/* ale_dr_getfluxvolume, ale_dr_getfluxvolume.f90:37-71 */
int orc_ale_getfluxvolume(orc_state_t* s, const bl_timestep_t* ts) {
  int nnd = s->sz.nnd;
  double *st4 = R(BL_RSCRATCH14), *st5 = R(BL_RSCRATCH15);
  if (s->cfg.zeul) { /* ale_kn_getmeshvelocity.f90:28-46 */
    for (int ind = 0; ind < nnd; ++ind) {
      st4[ind] = -st4[ind];
      st5[ind] = -st5[ind];
    }
  } else if (s->relax_type != BL_RELAX_OFF) {
    /* relaxation: w = (x_R - x)/dt from the Lagrangian positions */
    double* xr = (double*)malloc(sizeof(double) * (nnd > 0 ? nnd : 1));
    double* yr = (double*)malloc(sizeof(double) * (nnd > 0 ? nnd : 1));
    relax_sweeps(s, s->relax_type, s->relax_nsweep, s->relax_om, s->relax_tol, R(BL_NDX),
                 R(BL_NDY), xr, yr);
    const double *x = R(BL_NDX), *y = R(BL_NDY);
    for (int ind = 0; ind < nnd; ++ind) {
      st4[ind] = (xr[ind] - x[ind]) / ts->dt;
      st5[ind] = (yr[ind] - y[ind]) / ts->dt;
    }
    free(xr); free(yr);
  }
  k_copy(nnd, R(BL_NDX), R(BL_RSCRATCH12));
  k_copy(nnd, R(BL_NDY), R(BL_RSCRATCH13));
  k_getvertex(nnd, ts->dt, R(BL_NDX), R(BL_NDY), st4, st5);
  k_getfluxvolume(s->sz.nel, s->cfg.zerocut, I(BL_IELND), R(BL_RSCRATCH12),
                  R(BL_RSCRATCH13), R(BL_NDX), R(BL_NDY), R(BL_RSCRATCH21));
  return BL_SUCCESS;
}

/* ale_dr_advect*, ale_dr_advect.f90:56-305 (alestep aliases,
 * dataAPI_id.f90:102-111) */
static void ale_advectel(orc_state_t* s, int id1, int id2) {
  int nel = s->sz.nel;
  double *fcdv = R(BL_RSCRATCH21), *fcdm = R(BL_RSCRATCH22),
         *flux = R(BL_RSCRATCH23), *rwork1 = R(BL_RSCRATCH24),
         *rwork2 = R(BL_RSCRATCH25);
  /* ale_dr_advectbasisel */
  k_sumflux(id1, id2, nel, I(BL_IELEL), I(BL_IELFC), fcdv, rwork1);
  k_fluxelvl(id1, id2, s->sz.nel1, I(BL_IELEL), I(BL_IELFC), R(BL_CNWT), fcdv,
             R(BL_ELDENSITY), fcdm);
  k_sumflux(id1, id2, nel, I(BL_IELEL), I(BL_IELFC), fcdm, rwork2);
  k_updatebasisel(nel, s->cfg.zerocut, s->cfg.dencut, rwork1, rwork2,
                  R(BL_RSCRATCH15), R(BL_RSCRATCH16), R(BL_RSCRATCH11),
                  R(BL_RSCRATCH12), R(BL_RSCRATCH13), R(BL_ELVOLUME),
                  R(BL_ELMASS), R(BL_ELDENSITY));
  /* ale_dr_advectvarel */
  k_fluxelvl(id1, id2, s->sz.nel1, I(BL_IELEL), I(BL_IELFC), R(BL_CNMASS),
             fcdm, R(BL_ELENERGY), flux);
  k_updateel(id1, id2, nel, I(BL_IELEL), I(BL_IELFC), R(BL_RSCRATCH12),
             R(BL_ELMASS), R(BL_RSCRATCH16), flux, rwork1, R(BL_ELENERGY));
}

static void ale_advectnd(orc_state_t* s, int id1, int id2) {
  int nel = s->sz.nel, nnd = s->sz.nnd;
  double *fcdv = R(BL_RSCRATCH21), *fcdm = R(BL_RSCRATCH22),
         *flux = R(BL_RSCRATCH23), *rwork1 = R(BL_RSCRATCH24),
         *rwork2 = R(BL_RSCRATCH25), *rwork3 = R(BL_RSCRATCH26),
         *cnu = R(BL_RSCRATCH27), *cnv = R(BL_RSCRATCH28);
  double *st1 = R(BL_RSCRATCH11), *st2 = R(BL_RSCRATCH12),
         *st3 = R(BL_RSCRATCH13), *st4 = R(BL_RSCRATCH14),
         *st5 = R(BL_RSCRATCH15), *st6 = R(BL_RSCRATCH16);
  k_cngather(nel, I(BL_IELND), R(BL_NDU), cnu);
  k_cngather(nel, I(BL_IELND), R(BL_NDV), cnv);
  /* ale_dr_advectbasisnd */
  for (int ind = 0; ind < s->sz.nnd2; ++ind) { /* ale_kn_initbasisnd */
    st3[ind] = 0.0; st4[ind] = 0.0; st2[ind] = 0.0;
  }
  k_calcbasisnd(s->sz.nel2, I(BL_IELND), I(BL_IELSORT2), st1, R(BL_ELVOLUME),
                R(BL_CNMASS), st3, st4, st2, rwork3);
  k_fluxbasisnd(id1, id2, s->sz.nel2, I(BL_IELEL), I(BL_IELFC),
                I(BL_IELSORT2), fcdv, fcdm, rwork1, rwork2, flux);
  k_copy(s->sz.nnd2, st2, st1);
  k_massbasisnd(s->sz.nel2, I(BL_IELND), I(BL_IELSORT2), flux, R(BL_CNMASS),
                st1);
  for (int ind = 0; ind < nnd; ++ind) { /* ale_kn_cutbasisnd */
    st5[ind] = s->cfg.zerocut;
    st6[ind] = s->cfg.dencut * st3[ind];
  }
  /* ale_dr_advectvarnd */
  k_activend(nnd, -1, I(BL_ISCRATCH11), I(BL_INDTYPE), Z(BL_ZSCRATCH11));
  k_fluxndvl(s->sz.nel1, I(BL_IELEL), I(BL_IELFC), rwork3, rwork2, cnu, flux);
  k_updatend(nnd, s->sz.nel1, I(BL_IELND), st2, st1, st6, Z(BL_ZSCRATCH11),
             flux, fcdm, R(BL_NDU));
  k_activend(nnd, -2, I(BL_ISCRATCH11), I(BL_INDTYPE), Z(BL_ZSCRATCH11));
  k_fluxndvl(s->sz.nel1, I(BL_IELEL), I(BL_IELFC), rwork3, rwork2, cnv, flux);
  k_updatend(nnd, s->sz.nel1, I(BL_IELND), st2, st1, st6, Z(BL_ZSCRATCH11),
             flux, fcdm, R(BL_NDV));
}

int orc_ale_advect(orc_state_t* s, const bl_timestep_t* ts, bl_error_t* err) {
  switch (s->cfg.adv_type) {
    case 1:
      ale_advectel(s, 1, 2);
      ale_advectnd(s, 1, 2);
      break;
    case 2: {
      int ii = (ts->nstep + 1) % 2;
      int i1 = 1 + ii, i2 = 2 - ii, i3 = i2 - i1;
      for (ii = i1; (i3 > 0) ? (ii <= i2) : (ii >= i2); ii += i3) {
        ale_advectel(s, ii, ii);
        ale_advectnd(s, ii, ii);
      }
      break;
    }
    default:
      set_err(err, BL_FAILURE, BL_HALT_ALL, "ERROR: unrecognised adv_type");
      return BL_FAILURE;
  }
  set_err(err, BL_SUCCESS, 0, "");
  return BL_SUCCESS;
}

/* ale_dr_hydroupdate, ale_dr_update.f90:33-51 */
int orc_ale_hydroupdate(orc_state_t* s, bl_error_t* err) {
  int rc = orc_geometry_getgeometry(s, err);
  orc_density_getdensity(s);
  orc_hydro_geteos(s);
  return rc;
}

/* time_dr_calc + time_dr_end, time_dr_advance.f90:38-132 */
int orc_time_select(orc_state_t* s, bl_timestep_t* ts, const bl_dt_t* cfl,
                    const bl_dt_t* div, bl_error_t* err) {
  const bl_config_t* c = &s->cfg;
  bl_dt_t cand[5];
  int n = 0;
  cand[n].rdt = c->dt_max; cand[n].idt = -1; cand[n].sdt = BL_DT_MAXIMUM; ++n;
  cand[n].rdt = c->dt_g * ts->dt; cand[n].idt = -1; cand[n].sdt = BL_DT_GROWTH; ++n;
  if (ts->nstep == 0) {
    cand[n].rdt = c->dt_initial; cand[n].idt = -1; cand[n].sdt = BL_DT_INITIAL; ++n;
  }
  cand[n++] = *cfl;
  cand[n++] = *div;
  bl_dt_t mindt = cand[0];
  for (int i = 1; i < n; ++i)
    if (cand[i].rdt < mindt.rdt) mindt = cand[i];
  ts->nstep = ts->nstep + 1;
  ts->dt = mindt.rdt;
  ts->sdt = mindt.sdt;
  ts->idtel = mindt.idt;
  ts->time = ts->time + ts->dt;
  if (ts->dt < c->dt_min) {
    set_err(err, BL_FAILURE, BL_HALT_ALL, "ERROR: dt < dt_min");
    return BL_FAILURE;
  }
  set_err(err, BL_SUCCESS, 0, "");
  return BL_SUCCESS;
}

/* inf_solver_getdt + inf_solver_lagstep + inf_solver_alestep
 * (inf_solver_getdt.f90:35-70, inf_solver_lagstep.f90:44-129,
 * inf_solver_alestep.f90:39-74) */
int orc_step(orc_state_t* s, bl_timestep_t* ts, bl_error_t* err) {
  bl_dt_t cfl, div;
  if (orc_hydro_getdt(s, &cfl, &div, err) != BL_SUCCESS) return BL_FAILURE;
  if (orc_time_select(s, ts, &cfl, &div, err) != BL_SUCCESS) return BL_FAILURE;
  /* predictor */
  ts->zcorrector = 0;
  ts->dts = 0.5 * ts->dt;
  orc_hydro_setpredictor(s);
  orc_hydro_getforce(s, ts);
  if (orc_hydro_getstate(s, ts, err) != BL_SUCCESS) return BL_FAILURE;
  orc_hydro_getenergy(s, ts);
  orc_hydro_geteos(s);
  /* corrector */
  ts->zcorrector = 1;
  ts->dts = ts->dt;
  orc_hydro_setcorrector(s);
  orc_hydro_getforce(s, ts);
  orc_hydro_initacceleration(s);
  orc_hydro_scatteracceleration(s);
  orc_hydro_getacceleration(s);
  orc_hydro_applyacceleration(s, ts);
  if (orc_hydro_getstate(s, ts, err) != BL_SUCCESS) return BL_FAILURE;
  orc_hydro_getenergy(s, ts);
  orc_hydro_geteos(s);
  /* ALE */
  if (orc_ale_active(s, ts)) {
    orc_ale_getmeshstatus(s);
    orc_ale_getfluxvolume(s, ts);
    if (orc_ale_advect(s, ts, err) != BL_SUCCESS) return BL_FAILURE;
    if (orc_ale_hydroupdate(s, err) != BL_SUCCESS) return BL_FAILURE;
  }
  set_err(err, BL_SUCCESS, 0, "");
  return BL_SUCCESS;
}

int orc_run(orc_state_t* s, int32_t nsteps, int32_t stop_at_time_end,
            bl_timestep_t* ts, double* dt_hist, int32_t* idt_hist,
            int32_t* ctl_hist, int32_t* steps_taken, bl_error_t* err) {
  int taken = 0;
  int rc = BL_SUCCESS;
  set_err(err, BL_SUCCESS, 0, "");
  for (int k = 0; k < nsteps; ++k) {
    if (stop_at_time_end && ts->time >= s->cfg.time_end) break;
    rc = orc_step(s, ts, err);
    if (dt_hist) dt_hist[k] = ts->dt;
    if (idt_hist) idt_hist[k] = ts->idtel;
    if (ctl_hist) ctl_hist[k] = ts->sdt;
    ++taken;
    if (rc != BL_SUCCESS) break;
  }
  if (steps_taken) *steps_taken = taken;
  return rc;
}

/* totals in the sense of hydro_kn_calc_shortprint, hydro_kn_print.F90:56-110
 * (whole mesh as one flag; element order summation) */
int orc_totals(orc_state_t* s, bl_totals_t* out) {
  int nel = s->sz.nel;
  const double *den = R(BL_ELDENSITY), *en = R(BL_ELENERGY),
               *pre = R(BL_ELPRESSURE), *vol = R(BL_ELVOLUME),
               *mass = R(BL_ELMASS), *cnwt = R(BL_CNWT), *ndu = R(BL_NDU),
               *ndv = R(BL_NDV);
  const int32_t* ielnd = I(BL_IELND);
  double tv = 0.0, tm = 0.0, tie = 0.0, tke = 0.0;
  double pmn = DBL_MAX, pmx = -DBL_MAX, dmn = DBL_MAX, dmx = -DBL_MAX;
  for (int ii = 1; ii <= nel; ++ii) {
    double c1 = s->cfg.dencut * EL(vol, ii);
    tv = tv + EL(vol, ii);
    if (EL(mass, ii) > c1) {
      double w2 = EL(mass, ii);
      tm = tm + w2;
      double w3 = EL(en, ii) * w2;
      tie = tie + w3;
      double w4 = EL(pre, ii);
      if (w4 > pmx) pmx = w4;
      if (w4 < pmn) pmn = w4;
      w4 = EL(den, ii);
      if (w4 > dmx) dmx = w4;
      if (w4 < dmn) dmn = w4;
      for (int ic = 1; ic <= 4; ++ic) {
        int nd = CN(ielnd, ic, ii);
        double u = EL(ndu, nd), v = EL(ndv, nd);
        tke = 0.5 * CN(cnwt, ic, ii) * w4 * (u * u + v * v) + tke;
      }
    }
  }
  out->volume = tv; out->mass = tm; out->ie = tie; out->ke = tke;
  out->pressure_min = pmn; out->pressure_max = pmx;
  out->density_min = dmn; out->density_max = dmx;
  return BL_SUCCESS;
}

/* hydro_kn_init_shortprint + hydro_kn_calc_shortprint over one flag array
 * (hydro_kn_print.F90:32-110); cnu/cnv are the corners' velocities */
static void calc_shortprint(int nflag, int nsize, double dencut, const int32_t* flag,
                            const double* density, const double* energy,
                            const double* pressure, const double* volume,
                            const double* mass, const double* cnwt, const double* cnu,
                            const double* cnv, bl_flag_totals_t* t) {
  for (int ii = 1; ii <= nsize; ++ii) {
    int iflag = EL(flag, ii);
    if (iflag <= 0) continue;
    bl_flag_totals_t* f = &t[iflag - 1];
    double c1 = dencut * EL(volume, ii);
    f->volume = f->volume + EL(volume, ii);
    if (EL(mass, ii) > c1) {
      double w2 = EL(mass, ii);
      f->mass = f->mass + w2;
      double w3 = EL(energy, ii);
      w3 = w3 * w2;
      f->ie = f->ie + w3;
      double w4 = EL(pressure, ii);
      w3 = w2 * w4;
      f->pressure = f->pressure + w3;
      if (w4 > f->pressure_max) f->pressure_max = w4;
      if (w4 < f->pressure_min) f->pressure_min = w4;
      w4 = EL(density, ii);
      if (w4 > f->density_max) f->density_max = w4;
      if (w4 < f->density_min) f->density_min = w4;
      for (int ic = 1; ic <= 4; ++ic) {
        w2 = CN(cnu, ic, ii);
        w3 = CN(cnv, ic, ii);
        f->ke = 0.5 * CN(cnwt, ic, ii) * w4 * (w2 * w2 + w3 * w3) + f->ke;
      }
    }
  }
}

static void init_shortprint(int n, bl_flag_totals_t* t) {
  for (int k = 0; k < n; ++k) {
    t[k].volume = 0.0; t[k].ie = 0.0; t[k].pressure = 0.0;
    t[k].pressure_max = -DBL_MAX; t[k].pressure_min = DBL_MAX;
    t[k].mass = 0.0; t[k].ke = 0.0;
    t[k].density_max = -DBL_MAX; t[k].density_min = DBL_MAX;
  }
}

/* hydro_dr_shortprint, hydro_dr_print.f90:49-119 (tables only) */
int orc_shortprint(orc_state_t* s, bl_flag_totals_t* region, bl_flag_totals_t* material) {
  int nel = s->sz.nel, ncp = s->sz.ncp;
  double* cnu = (double*)malloc(sizeof(double) * 4 * (size_t)(nel > 0 ? nel : 1));
  double* cnv = (double*)malloc(sizeof(double) * 4 * (size_t)(nel > 0 ? nel : 1));
  k_cngather(nel, I(BL_IELND), R(BL_NDU), cnu);
  k_cngather(nel, I(BL_IELND), R(BL_NDV), cnv);
  init_shortprint(s->sz.nreg, region);
  calc_shortprint(s->sz.nreg, nel, s->cfg.dencut, I(BL_IELREG), R(BL_ELDENSITY),
                  R(BL_ELENERGY), R(BL_ELPRESSURE), R(BL_ELVOLUME), R(BL_ELMASS),
                  R(BL_CNWT), cnu, cnv, region);
  init_shortprint(s->sz.nmat, material);
  calc_shortprint(s->sz.nmat, nel, s->cfg.dencut, I(BL_IELMAT), R(BL_ELDENSITY),
                  R(BL_ELENERGY), R(BL_ELPRESSURE), R(BL_ELVOLUME), R(BL_ELMASS),
                  R(BL_CNWT), cnu, cnv, material);
  if (ncp > 0) { /* utils_kn_mxgathercncp of cnwt, cnu, cnv to the components */
    double* cw = (double*)malloc(sizeof(double) * 4 * (size_t)ncp);
    double* cu = (double*)malloc(sizeof(double) * 4 * (size_t)ncp);
    double* cv = (double*)malloc(sizeof(double) * 4 * (size_t)ncp);
    for (int imx = 1; imx <= s->sz.nmx; ++imx) {
      int iel = EL(I(BL_IMXEL), imx), icp = EL(I(BL_IMXFCP), imx);
      for (int ii = 1; ii <= EL(I(BL_IMXNCP), imx); ++ii, ++icp)
        for (int j = 1; j <= 4; ++j) {
          CN(cw, j, icp) = CN(R(BL_CNWT), j, iel);
          CN(cu, j, icp) = CN(cnu, j, iel);
          CN(cv, j, icp) = CN(cnv, j, iel);
        }
    }
    calc_shortprint(s->sz.nmat, ncp, s->cfg.dencut, I(BL_ICPMAT), R(BL_CPDENSITY),
                    R(BL_CPENERGY), R(BL_CPPRESSURE), R(BL_CPVOLUME), R(BL_CPMASS), cw, cu,
                    cv, material);
    free(cw); free(cu); free(cv);
  }
  free(cnu); free(cnv);
  return BL_SUCCESS;
}

/* check_kn_sod + check_dr_tests (check_kn_tests.f90:33-62,
 * check_dr_tests.f90:56-69) with the solution sod(scale*(xc-x0),
 * scale*time); centroid of geometry_kn_getcentroid
 * (geometry_kn_centroid.f90:29-55) */
int orc_sod_l1(orc_state_t* s, double time, double x0, double scale, double* l1) {
  int nel = s->sz.nel;
  const double *cnx = R(BL_CNX), *cny = R(BL_CNY);
  double basis = 0.0, la = 0.0, lb = 0.0;
  for (int ii = 1; ii <= nel; ++ii) {
    double cx = 0.0, cy = 0.0, rvol = 0.0;
    for (int i1 = 1; i1 <= 4; ++i1) {
      int i2 = (i1 % 4) + 1;
      double rr = CN(cnx, i1, ii) * CN(cny, i2, ii) - CN(cnx, i2, ii) * CN(cny, i1, ii);
      rvol = rvol + rr;
      cx = cx + rr * (CN(cnx, i1, ii) + CN(cnx, i2, ii));
      cy = cy + rr * (CN(cny, i1, ii) + CN(cny, i2, ii));
    }
    rvol = 0.5 * rvol;
    cx = (rvol > 0.0) ? cx / (6.0 * rvol) : 0.0;
    double r, e;
    orc_sod_exact(scale * (cx - x0), scale * time, &r, &e);
    double v = EL(R(BL_ELVOLUME), ii);
    basis = basis + v;
    la = fabs(EL(R(BL_ELDENSITY), ii) - r) * v + la;
    lb = fabs(EL(R(BL_ELENERGY), ii) - e) * v + lb;
  }
  l1[0] = la / basis;
  l1[1] = lb / basis;
  return BL_SUCCESS;
}

/* sod(), check_kn_tests.f90:64-111 */
void orc_sod_exact(double xx, double time, double* rho, double* e) {
  double coeff = sqrt(1.4);
  double r1 = -DBL_MAX, r2 = -DBL_MAX;
  if (xx <= -coeff * time) {
    r1 = 1.0; r2 = 2.5;
  } else if (xx <= -0.07027281256118326961 * time) {
    double ss = xx / time;
    double velocity = (5.0 / 6.0) * (ss + coeff);
    double tt = ((velocity - ss) * (velocity - ss)) / 1.4;
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
  *rho = r1;
  *e = r2;
}
