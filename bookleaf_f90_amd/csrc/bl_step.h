// bl_step.h -- host-side dt selection shared by the step runners.
#pragma once
#include "bl_internal.h"

namespace bl {

// time_dr_calc + time_dr_end (time_dr_advance.f90:38-132): strict '<' minimum
// over MAXIMUM, GROWTH, [INITIAL], CFL, DIV in that order
inline int select_dt(const bl_config_t& c, bl_timestep_t* ts, const bl_dt_t& cfl,
                     const bl_dt_t& div) {
  bl_dt_t cand[5];
  int n = 0;
  cand[n++] = {c.dt_max, -1, BL_DT_MAXIMUM};
  cand[n++] = {c.dt_g * ts->dt, -1, BL_DT_GROWTH};
  if (ts->nstep == 0) cand[n++] = {c.dt_initial, -1, BL_DT_INITIAL};
  cand[n++] = cfl;
  cand[n++] = div;
  bl_dt_t m = cand[0];
  for (int i = 1; i < n; ++i)
    if (cand[i].rdt < m.rdt) m = cand[i];
  ts->nstep = ts->nstep + 1;
  ts->dt = m.rdt;
  ts->sdt = m.sdt;
  ts->idtel = m.idt;
  ts->time = ts->time + ts->dt;
  return (ts->dt < c.dt_min) ? BL_FAILURE : BL_SUCCESS;
}

}  // namespace bl
