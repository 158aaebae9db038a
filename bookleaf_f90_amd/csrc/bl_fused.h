// bl_fused.h -- kernel arguments of the device-resident step (bl_fused.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bl_internal.h"
#include "bl_mix.h"

namespace bl {

// The run call's counters (bl_run_stats): the elements the limiter pass
// listed at vstats[0], and the count of elements whose zero viscosity a
// step wrote spread over kStatSlots counters one 128-byte line apart
// (vstats[kStatStride * (1 + s)]): a wave adds to its workgroup's slot, so
// that a step where every wave writes zeros does not queue all its atomics
// on one address.
constexpr int kStatSlots = 64;
constexpr int kStatStride = 16;
constexpr int kStatWords = kStatStride * (1 + kStatSlots);

struct FusedNodes {
  const double* x;
  const double* y;
  const double* u;
  const double* v;
};
struct FusedNodesOut {
  double* x;
  double* y;
  double* u;
  double* v;
};

// everything a step kernel needs, passed by value (kernel argument space)
struct FusedArgs {
  int nel, nnd;
  int reg_tables, reg_tables_force, zsp, zhg, zexist, write_geometry;
  double zcut, zerocut, dencut, accut2, cvisc1, cvisc2, cfl_sf, div_sf, ccut, pcut;
  double dt_g, dt_min, dt_max, dt_initial, mintime, maxtime;
  double time_end;        // with stop_at_time_end: steps from time_end on do nothing
  int stop_at_time_end;
  // connectivity
  const int32_t* ielnd;
  const int32_t* ielel;
  const int32_t* ielreg;
  const int32_t* ielmat;
  const int32_t* indtype;
  const int32_t* hyd_ptr;
  const int32_t* hyd_idx;
  const int32_t* outer;  // limiter stencil (8 far nodes per element) or null
  // tables
  const int32_t* zdtnotreg;
  const int32_t* zmidlength;
  const double* kappareg;
  const double* pmeritreg;
  const int32_t* eos_itype;
  const double* eos_param;
  // element fields (rho, en, pre, cs2 as the step starts; the corrector
  // writes rho_o, en_o, pre_o, cs2_o: the same arrays, or the other element
  // set in the tile step)
  double* rho;
  double* en;
  double* pre;
  double* cs2;
  double* rho_o;
  double* en_o;
  double* pre_o;
  double* cs2_o;
  double* vol;
  const double* mass;
  const double* cnmass;
  const double* spmass;
  double* cnwt;
  double* cnx;
  double* cny;
  double* a1;
  double* a2;
  double* a3;
  double* b1;
  double* b2;
  double* b3;
  double* cnviscx;
  double* cnviscy;
  double* elvisc;
  double* cnfx;  // rscratch23 (lagstep cnfx)
  double* cnfy;  // rscratch24
  // corrector corner terms in corner-major layout X[j*nel + e] (j = corner):
  // force x/y, half-step corner weight and the nodal-mass term with the
  // fallbacks of hydro_kn_scatteracceleration already resolved
  double* fxT;
  double* fyT;
  double* wtT;
  double* mT;
  double* ubar;  // rscratch14 (lagstep ndubar)
  double* vbar;  // rscratch15
  // reductions and device timestep
  double* part;
  int32_t* parti;
  // an ensemble's worklist of elements whose nodes move unequally, one
  // segment per instance (a single problem's F1L reads the at-rest mask)
  int32_t* work;
  // bit e of rest[e / 64]: element e is at rest this step and its cnviscx/y
  // and elvisc hold the +0.0 F1 wrote (F1 sets it; F2 reads it)
  uint64_t* rest;
  unsigned long long* vstats;  // kStatWords counters (kStatSlots above)
  // F1 folded into F4 (a Lagrangian step with pure cells): F4 does F1's work
  // for the next step at x^{n+1}, u^{n+1} and writes rest_next; F2 writes
  // the zeros of elements at rest now (rest) but not in the step before
  // (rest_prev).  Without the fold, F1 writes them and rest is its only mask.
  int fold;
  // no F5: the next step's F1b records this step's end (step_end), and the
  // host the last step's (a pure-cell Lagrangian batch without per-step
  // diagnostics)
  int end_in_f1b;
  const uint64_t* rest_prev;
  uint64_t* rest_next;
  DtResult* dtres;
  DevTimestep* ts;
  double* hist;
  int hist_cap;
  MixArgs mx;  // mixed cells (mx.ncp == 0: none)
  // the tile step (bl_tiles.h): tiles, their nodes (0-based | the tile
  // writes it << 31), their element entries (element | flags << 30, four
  // 8-bit local node slots) and the at-rest bytes of the step (rest8: now,
  // rest8_prev: the step before, rest8_next: written for the next step);
  // one F1 partial slot per tile (nseg = ntile)
  int ntile, nseg;
  const int32_t* tile_eptr;
  const int32_t* tile_nptr;
  const int4* tile_bounds;  // (tile_eptr[t], tile_eptr[t+1], tile_nptr[t], tile_nptr[t+1])
  const int32_t* tile_nodes;
  const uint8_t* tile_ncm;  // per tile node: bit jj, the node is corner jj of an element
  const int2* tile_el;
  const int4* tile_nd;  // the entry's four nodes (0-based), beside tile_el
  const uint8_t* rest8;
  const uint8_t* rest8_prev;
  uint8_t* rest8_next;
  // ensembles (null otherwise): the instance of each element and node, and
  // each instance's dt, which F2-F4 use in place of the timestep's
  const uint16_t* einst;
  const uint16_t* ninst;
  const double* inst_dt;
  const double* inst_rdt;  // 1/inst_dt, formed where each instance's dt is selected
};

// the timestep of one ensemble instance
struct InstTs {
  double time, dt;
  int32_t nstep, idtel, sdt, pad;
};

// ensemble step state: instance element offsets (ninst + 1), per-instance
// timesteps and worklist lengths, F1's two partial slots per workgroup
// (DIV a, DIV b, CFL a, CFL b, each nb long: a workgroup spans at most two
// instances) and F1L's partials (one per F1L workgroup), and the dt
// history, (step, instance, 3)
struct EnsArgs {
  int ninst;
  const int32_t* eoff;
  InstTs* inst;
  double* inst_dt;
  double* inst_rdt;
  uint32_t* cnt;
  double* part;
  int32_t* parti;
  const int32_t* lboff;  // F1L workgroups of instance k: lboff[k] .. lboff[k+1]-1
  double* lpart;
  int32_t* lparti;
  double* hist;
  int hist_cap;
  // with the tile step: instance k's tiles are toff[k] .. toff[k+1]-1, and
  // their F1 partials the tile step's slots (null: the element blocks')
  const int32_t* toff;
};

}  // namespace bl
