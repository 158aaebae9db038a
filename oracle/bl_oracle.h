/*
 * bl_oracle.h -- CPU restatement of BookLeaf's Lagrangian step and ALE remap.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP implementation
 * is compared against; it is loaded only by tests/, by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py.  Nothing in the product links it.
 *
 * Every function follows the cited reference routine statement by statement,
 * with the same left-to-right association, so that results are bit-identical
 * to the reference built without FMA contraction.  Arrays are the raw dataAPI
 * arrays (Fortran column-major, (NCORN,nel) corner arrays, 1-based index
 * contents).
 */
#ifndef BL_ORACLE_H
#define BL_ORACLE_H

#include <stdint.h>
#include "../include/bookleaf_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state_t;

orc_state_t* orc_create(const bl_sizes_t* sizes);
void orc_destroy(orc_state_t* s);
void orc_set_config(orc_state_t* s, const bl_config_t* cfg);
void orc_set_regions(orc_state_t* s, int32_t nreg, const double* kappareg,
                     const double* pmeritreg, const int32_t* zdtnotreg,
                     const int32_t* zmidlength);
void orc_set_eos(orc_state_t* s, int32_t nmat, const int32_t* itype,
                 const double* param);
void* orc_field(orc_state_t* s, int32_t id, int32_t* kind, int64_t* count);
void orc_get_sizes(orc_state_t* s, bl_sizes_t* out);

/* drivers, same meaning as the bl_* entry points of bookleaf_hip.h */
int orc_init_mesh(orc_state_t* s);
int orc_init_state(orc_state_t* s);
int orc_hydro_init(orc_state_t* s);
int orc_hydro_getdt(orc_state_t* s, bl_dt_t* cfl, bl_dt_t* div,
                    bl_error_t* err);
int orc_hydro_setpredictor(orc_state_t* s);
int orc_hydro_setcorrector(orc_state_t* s);
int orc_hydro_getforce(orc_state_t* s, const bl_timestep_t* ts);
int orc_hydro_getstate(orc_state_t* s, const bl_timestep_t* ts,
                       bl_error_t* err);
int orc_hydro_getenergy(orc_state_t* s, const bl_timestep_t* ts);
int orc_hydro_geteos(orc_state_t* s);
int orc_hydro_initacceleration(orc_state_t* s);
int orc_hydro_scatteracceleration(orc_state_t* s);
int orc_hydro_getacceleration(orc_state_t* s);
int orc_hydro_applyacceleration(orc_state_t* s, const bl_timestep_t* ts);
int orc_geometry_getgeometry(orc_state_t* s, bl_error_t* err);
int orc_density_getdensity(orc_state_t* s);
int orc_ale_active(orc_state_t* s, const bl_timestep_t* ts);
int orc_ale_getmeshstatus(orc_state_t* s);
int orc_ale_getfluxvolume(orc_state_t* s, const bl_timestep_t* ts);
int orc_ale_advect(orc_state_t* s, const bl_timestep_t* ts, bl_error_t* err);
int orc_ale_hydroupdate(orc_state_t* s, bl_error_t* err);

/* time_dr_calc + time_dr_end (time_dr_advance.f90:38-132) on host scalars:
 * combines the hydro candidates with MAXIMUM/GROWTH/INITIAL, advances ts */
int orc_time_select(orc_state_t* s, bl_timestep_t* ts, const bl_dt_t* cfl,
                    const bl_dt_t* div, bl_error_t* err);
/* one full step: inf_solver_getdt + lagstep + alestep */
int orc_step(orc_state_t* s, bl_timestep_t* ts, bl_error_t* err);
int orc_run(orc_state_t* s, int32_t nsteps, int32_t stop_at_time_end,
            bl_timestep_t* ts, double* dt_hist, int32_t* idt_hist,
            int32_t* ctl_hist, int32_t* steps_taken, bl_error_t* err);
int orc_totals(orc_state_t* s, bl_totals_t* out);
int orc_shortprint(orc_state_t* s, bl_flag_totals_t* region, bl_flag_totals_t* material);
int orc_sod_l1(orc_state_t* s, double time, double x0, double scale, double* l1);

/* node relaxation (bl_set_relaxation / bl_relax_mesh of the C-ABI) */
int orc_set_relaxation(orc_state_t* s, int32_t type, int32_t nsweep, double om,
                       double tol);
int orc_relax_mesh(orc_state_t* s, int32_t type, int32_t nsweep, double om, double tol);
const int32_t* orc_relax_ring(orc_state_t* s);
void orc_relax_plant(orc_state_t* s, int32_t which);

/* analytic Sod solution of check_kn_tests.f90:64-111 at (x,t) for the
 * reference's [0,100] tube; returns density and energy */
void orc_sod_exact(double x, double t, double* rho, double* e);

#ifdef __cplusplus
}
#endif
#endif
