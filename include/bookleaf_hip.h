/*
 * bookleaf_hip.h -- C-ABI of the MI355X (gfx950) implementation of BookLeaf's
 * Lagrangian step and ALE remap.
 *
 * The library keeps a device-resident mirror of the dataAPI field registry
 * (ids and shapes of src/dataAPI/dataAPI_id.f90:26-111 and
 * src/utilities/data/data.f90:169-430) and exposes one entry point per
 * reference driver on the hot path, so a Fortran host can replace the bodies
 * of hydro_dr_* / ale_dr_* / geometry_dr_* / eos_dr_* / density_dr_* with
 * iso_c_binding calls into this library and link the rest of the reference
 * unchanged.  All arguments are plain C types: int32 sizes, FP64 scalars and
 * raw host pointers.  Integer index fields (ielnd, ielel, ielfc, ielsort*)
 * hold the reference's 1-based values unchanged.
 *
 * Conventions
 *   - every function returns BL_SUCCESS (0) or BL_FAILURE (-1)
 *     (dataAPI_params.f90:31-34); where an error_t is passed it is filled like
 *     the reference fills it (ierr, iout = BL_HALT_ALL/BL_HALT_SINGLE, serr).
 *   - calls are stream-ordered on the handle's own HIP stream; calls that
 *     return a host scalar (getdt, getgeometry's negative-volume check,
 *     downloads) synchronise that stream first.
 *   - one partition per GPU: nel1=nel2=nel, nnd1=nnd2=nnd, and every Typhon
 *     exchange of the reference is a no-op (nproc=1).
 */
#ifndef BOOKLEAF_HIP_H
#define BOOKLEAF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- dataAPI_params.f90:25-34 ---------------------------------------- */
#define BL_NCORN 4
#define BL_NFACE 4
#define BL_SUCCESS 0
#define BL_FAILURE (-1)
#define BL_HALT_ALL 1
#define BL_HALT_SINGLE 2
#define BL_SLEN 130
#define BL_NEOSPARAM 6 /* setup_read.f90:518 eos_param(6,LN) */

/* ---- field ids: dataAPI_id.f90:26-74 (1-based, identical numbering) --- */
enum bl_field_id {
  BL_ELDENSITY = 1, BL_ELENERGY = 2, BL_ELPRESSURE = 3, BL_ELCS2 = 4,
  BL_ELVOLUME = 5, BL_A1 = 6, BL_A2 = 7, BL_A3 = 8, BL_B1 = 9, BL_B2 = 10,
  BL_B3 = 11, BL_CNWT = 12, BL_CNX = 13, BL_CNY = 14,
  BL_RSCRATCH21 = 15, BL_RSCRATCH22 = 16, BL_RSCRATCH23 = 17,
  BL_RSCRATCH24 = 18, BL_RSCRATCH25 = 19, BL_RSCRATCH26 = 20,
  BL_RSCRATCH27 = 21, BL_RSCRATCH28 = 22,
  BL_CNVISCX = 23, BL_CNVISCY = 24, BL_ELVISC = 25, BL_ELMASS = 26,
  BL_CNMASS = 27, BL_RSCRATCH11 = 28, BL_RSCRATCH12 = 29, BL_RSCRATCH13 = 30,
  BL_RSCRATCH14 = 31, BL_RSCRATCH15 = 32, BL_RSCRATCH16 = 33,
  BL_RSCRATCH17 = 34, BL_NDX = 35, BL_NDY = 36, BL_NDU = 37, BL_NDV = 38,
  BL_INDTYPE = 39, BL_IELSORT1 = 40, BL_IELND = 41, BL_IELEL = 42,
  BL_IELFC = 43, BL_IELMAT = 44, BL_IELREG = 45, BL_IELSORT2 = 46,
  BL_SPMASS = 47, BL_ISCRATCH11 = 48, BL_ZSCRATCH11 = 49,
  BL_IELLOCGLOB = 50, BL_INDLOCGLOB = 51,
  BL_CPDENSITY = 52, BL_CPENERGY = 53, BL_CPPRESSURE = 54, BL_CPCS2 = 55,
  BL_CPVOLUME = 56, BL_FRVOLUME = 57, BL_CPMASS = 58, BL_FRMASS = 59,
  BL_CPVISCX = 60, BL_CPVISCY = 61, BL_CPVISC = 62, BL_CPA1 = 63,
  BL_CPA3 = 64, BL_CPB1 = 65, BL_CPB3 = 66, BL_ICPMAT = 67, BL_ICPNEXT = 68,
  BL_ICPPREV = 69, BL_IMXEL = 70, BL_IMXFCP = 71, BL_IMXNCP = 72,
  BL_ICPSCRATCH11 = 73, BL_ICPSCRATCH12 = 74, BL_RCPSCRATCH11 = 75,
  BL_RCPSCRATCH21 = 76, BL_RCPSCRATCH22 = 77, BL_RCPSCRATCH23 = 78,
  BL_RCPSCRATCH24 = 79,
  BL_NFIELD = 79
};

/* element type of a field, as registered by data.f90 (rlk / ink / lak) */
enum bl_field_kind { BL_KIND_R64 = 1, BL_KIND_I32 = 2, BL_KIND_BOOL = 3 };

/* dt control names of time_dr_advance.f90:38-72 and hydro_kn_getdt.f90:82,121 */
enum bl_dt_control {
  BL_DT_MAXIMUM = 1, BL_DT_GROWTH = 2, BL_DT_INITIAL = 3, BL_DT_CFL = 4,
  BL_DT_DIV = 5
};

/* sizes_t, dataAPI_types.f90:41-44 */
typedef struct bl_sizes {
  int32_t nel, nel1, nel2, nnd, nnd1, nnd2, nsz, nmat, nreg, ncp, nmx, mcp,
      mmx;
} bl_sizes_t;

/* Scalar configuration.  Fields mirror global_t (global_types.f90:25-27),
 * hydro_t (hydro_types.f90:28-39), eos_t (eos_types.f90), time_t
 * (time_types.f90) and ale_t (ale_types.f90:27-42).  Logical flags are int32
 * 0/1.  Per-region and per-material tables are passed by bl_set_regions and
 * bl_set_eos. */
typedef struct bl_config {
  double zcut, zerocut, dencut, accut;          /* global_t */
  double cvisc1, cvisc2, cfl_sf, div_sf;        /* hydro_t  */
  int32_t zhg, zsp, ztq, pad0;                  /* hydro_t  */
  double ccut, pcut;                            /* eos_t    */
  double time_start, time_end, dt_g, dt_min, dt_max, dt_initial; /* time_t */
  double mintime, maxtime, sf;                  /* ale_t    */
  int32_t adv_type, zexist, zon, zeul;          /* ale_t    */
} bl_config_t;

/* timestep_t, dataAPI_types.f90:46-52 (sdt/mdt strings are replaced by the
 * bl_dt_control code; the host derives mdt from ielmat(idtel)) */
typedef struct bl_timestep {
  double time, dt, dts;
  int32_t nstep, idtel, zcorrector, sdt;
} bl_timestep_t;

/* dt_t, dataAPI_types.f90:54-60 (without the linked-list pointer) */
typedef struct bl_dt {
  double rdt;
  int32_t idt, sdt;
} bl_dt_t;

/* error_t, dataAPI_types.f90:76-79 */
typedef struct bl_error {
  int32_t ierr, iout;
  char serr[BL_SLEN];
} bl_error_t;

/* per-step conservation totals in the sense of hydro_kn_calc_shortprint
 * (hydro_kn_print.F90:56-110), summed over the whole mesh */
typedef struct bl_totals {
  double volume, mass, ie, ke, pressure_min, pressure_max, density_min,
      density_max;
} bl_totals_t;

/* one row of hydro_kn_calc_shortprint's tables (hydro_kn_print.F90:56-110)
 * for a region or a material: pressure is the mass-weighted sum */
typedef struct bl_flag_totals {
  double volume, mass, ie, ke, pressure, pressure_min, pressure_max,
      density_min, density_max;
} bl_flag_totals_t;

typedef struct bl_handle bl_handle_t;

/* ---- lifetime (data.f90:129-167 data_set / data_setinit) -------------- */
/* Allocate the device mirror of every dataAPI field for these sizes on the
 * current HIP device.  Real fields are filled with RINIT=-2e12, integer
 * fields with IINIT=-2e9 and logical fields with .FALSE. (data.f90:107-109),
 * as the reference registry does.  sizes->nsz is recomputed as
 * MAX(nel2,nnd2) (data.f90:165). */
int bl_create(const bl_sizes_t* sizes, bl_handle_t** out);
int bl_destroy(bl_handle_t* h);
int bl_set_config(bl_handle_t* h, const bl_config_t* cfg);
/* hydro_t per-region tables: kappareg, pmeritreg, zdtnotreg, zmidlength
 * (hydro_types.f90:32-34); length nreg each, logicals as int32. */
int bl_set_regions(bl_handle_t* h, int32_t nreg, const double* kappareg,
                   const double* pmeritreg, const int32_t* zdtnotreg,
                   const int32_t* zmidlength);
/* eos_t tables after eos_cf_rationalise (eos_config.f90:163-226): itype and
 * param for materials 0..nmat, param stored (BL_NEOSPARAM, 0:nmat). */
int bl_set_eos(bl_handle_t* h, int32_t nmat, const int32_t* itype,
               const double* param);

/* ---- field transfer (the dh(id)%raddr / %iaddr / %zaddr arrays) ------- */
int bl_field_info(bl_handle_t* h, int32_t id, int32_t* kind, int64_t* count);
void* bl_field_device_ptr(bl_handle_t* h, int32_t id);
int bl_upload(bl_handle_t* h, int32_t id, const void* host, int64_t count);
int bl_download(bl_handle_t* h, int32_t id, void* host, int64_t count);
/* Build the node->corner CSR maps from ielnd, ielsort1, ielsort2 (upload
 * those first).  Two orderings: hydro (corner jj, then element in ielsort1
 * order, hydro_kn_getacceleration.f90:53-72) and ALE (element in ielsort2
 * order, then corner, ale_kn_advect.f90:96-124). */
int bl_build_maps(bl_handle_t* h);
/* Host-side mesh initialisation for hosts that do not run the reference's
 * own init: ielel/ielfc from ielnd (init_kn_connectivity, init_kn.f90:82-91,
 * utils_kn_connectivity.f90:30-170), corner node types (init_kn_nodetype,
 * init_kn.f90:93-113), ielsort1/2 = identity (init_dr_utils.f90:64-69), then
 * bl_build_maps.  Upload ielnd and indtype first. */
int bl_init_mesh(bl_handle_t* h);
int bl_synchronize(bl_handle_t* h);

/* ---- initialisation drivers ------------------------------------------ */
/* init_dr_state: elmass=rho*V, cnmass=rho*cnwt, then EOS
 * (init_dr_utils.f90:93-117; init_kn.f90:32-48) */
int bl_init_state(bl_handle_t* h);
/* hydro_dr_init: zero viscosity, sub-zonal masses if zsp
 * (hydro_dr_init.f90:37-57) */
int bl_hydro_init(bl_handle_t* h);

/* ---- Lagrangian step drivers ----------------------------------------- */
/* hydro_dr_getdt (hydro_dr_getdt.f90:39-79): corner gather of ndu/ndv,
 * edge viscosity, cs2 += visc, CFL and divergence controls. */
int bl_hydro_getdt(bl_handle_t* h, bl_dt_t* cfl, bl_dt_t* div,
                   bl_error_t* err);
/* bl_hydro_getdt without the handle's reduction buffers: no counterpart in
 * the reference.  For callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc): the call allocates and frees
 * nothing.  Same kernels, same results and the same fields written as
 * bl_hydro_getdt; its partial minima/maxima and their result live in a
 * library-wide device array (one per GPU, about 2.3 MiB, held until the
 * process ends) whose address bl_create looks up.  Limit: a mesh of at most
 * BL_NOALLOC_MAX_EL elements, else BL_FAILURE with the reason in
 * bl_last_error; the handle's device must be the current device.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (two host
 * threads, two handles' streams) race and must be serialised by the
 * caller. */
#define BL_NOALLOC_MAX_EL (1 << 24)
int bl_hydro_getdt_noalloc(bl_handle_t* h, bl_dt_t* cfl, bl_dt_t* div,
                           bl_error_t* err);
/* hydro_dr_setpredictor / setcorrector (hydro_dr_set.f90:31-51) */
int bl_hydro_setpredictor(bl_handle_t* h);
int bl_hydro_setcorrector(bl_handle_t* h);
/* hydro_dr_getforce (hydro_dr_getforce.f90:40-91) */
int bl_hydro_getforce(bl_handle_t* h, const bl_timestep_t* ts);
/* hydro_dr_getstate (hydro_dr_getstate.f90:33-51): vertex update with
 * 0.5*dt, geometry, density */
int bl_hydro_getstate(bl_handle_t* h, const bl_timestep_t* ts,
                      bl_error_t* err);
/* hydro_dr_getenergy (hydro_dr_getenergy.f90:35-64) */
int bl_hydro_getenergy(bl_handle_t* h, const bl_timestep_t* ts);
/* hydro_dr_geteos / eos_dr_geteos (eos_dr_geteos.f90:35-66) */
int bl_hydro_geteos(bl_handle_t* h);
/* hydro_dr_getacceleration.f90:44-112, split as the reference splits it */
int bl_hydro_initacceleration(bl_handle_t* h);
int bl_hydro_scatteracceleration(bl_handle_t* h);
int bl_hydro_getacceleration(bl_handle_t* h);
int bl_hydro_applyacceleration(bl_handle_t* h, const bl_timestep_t* ts);

/* ---- shared utilities ------------------------------------------------- */
/* geometry_dr_getgeometry (geometry_dr_volume.f90:40-96); on a negative
 * volume err->ierr = -iel of the LAST such element, iout=HALT_SINGLE */
int bl_geometry_getgeometry(bl_handle_t* h, bl_error_t* err);
/* density_dr_getdensity (density_dr_getdensity.f90:32-46) */
int bl_density_getdensity(bl_handle_t* h);

/* ---- ALE step drivers ------------------------------------------------- */
/* ale_active (ale_dr_utils.f90:35-51) */
int bl_ale_active(bl_handle_t* h, const bl_timestep_t* ts);
int bl_ale_getmeshstatus(bl_handle_t* h);                  /* ale_dr_getmeshstatus.f90:32-49 */
int bl_ale_getfluxvolume(bl_handle_t* h, const bl_timestep_t* ts); /* ale_dr_getfluxvolume.f90:37-71 */
int bl_ale_advect(bl_handle_t* h, const bl_timestep_t* ts,
                  bl_error_t* err);                         /* ale_dr_advect.f90:56-100 */
int bl_ale_hydroupdate(bl_handle_t* h, bl_error_t* err);   /* ale_dr_update.f90:33-51 */

/* ---- whole-step residency --------------------------------------------- */
/* Run up to nsteps complete steps (inf_solver_getdt + inf_solver_lagstep +
 * inf_solver_alestep, inf_control_solver.f90:53-69) with the dt selection of
 * time_dr_calc/time_dr_end (time_dr_advance.f90:38-132) on the device and no
 * host synchronisation inside the batch.  ts is read at entry and written at
 * exit.  If stop_at_time_end != 0 the steps after the one that reaches
 * time_end (time_dr_utils.f90:57-69) do nothing.  Optional history arrays
 * (length nsteps, may be NULL) receive dt, idtel and the bl_dt_control code
 * of every step taken.  *steps_taken receives the number of steps taken.
 * Errors (negative volume, negative CFL, dt < dt_min) stop the batch at the
 * failing step and are reported through err exactly as the host drivers
 * would report them. */
int bl_run(bl_handle_t* h, int32_t nsteps, int32_t stop_at_time_end,
           bl_timestep_t* ts, double* dt_hist, int32_t* idt_hist,
           int32_t* ctl_hist, int32_t* steps_taken, bl_error_t* err);
/* Same as bl_run, replaying one captured hipGraph per step (identical
 * kernels and arguments; dt never leaves the device). */
int bl_run_graph(bl_handle_t* h, int32_t nsteps, int32_t stop_at_time_end,
                 bl_timestep_t* ts, double* dt_hist, int32_t* idt_hist,
                 int32_t* ctl_hist, int32_t* steps_taken, bl_error_t* err);

/* Conservation totals over all elements (hydro_kn_print.F90:56-110):
 * volume, mass of elements with m > dencut*V, IE = sum m*e,
 * KE = sum_c 0.5*rho*cnwt_c*|u_c|^2 with u_c gathered from ndu/ndv. */
int bl_totals(bl_handle_t* h, bl_totals_t* out);

/* hydro_dr_shortprint's tables (hydro_dr_print.f90:49-119): region[nreg]
 * over elements by ielreg, material[nmat] over pure elements by ielmat plus
 * the components of mixed cells by icpmat (component corners take the
 * cell's cnwt and corner velocities, utils_kn_mxgathercncp).  Sums use a
 * fixed two-level tree (per workgroup, then over workgroups in order), so
 * results are reproducible run to run; they differ from the serial sums in
 * the last bits. */
int bl_shortprint(bl_handle_t* h, bl_flag_totals_t* region, bl_flag_totals_t* material);
/* check_kn_sod (check_kn_tests.f90:33-62, check_dr_tests.f90:56-69): volume
 * weighted L1 of density and energy against sod(scale*(xc - x0),
 * scale*time) at the element centroids; l1[0] density, l1[1] energy.  The
 * reference's own check is x0 = 50, scale = 1. */
int bl_sod_l1(bl_handle_t* h, double time, double x0, double scale, double* l1);
/* Per-step diagnostics of bl_run_graph: with on != 0 every step records
 * nstep, time, the region table (nreg rows of bl_flag_totals_t) and the Sod
 * L1 pair for (x0, scale) on the device, inside the captured step.
 * bl_diagnostics copies up to maxrec records of 2 + 9*nreg + 2 doubles each
 * (oldest first, since the last bl_set_diagnostics) and sets *nrec. */
int bl_set_diagnostics(bl_handle_t* h, int32_t on, double x0, double scale, int32_t maxrec);
int bl_diagnostics(bl_handle_t* h, int32_t maxrec, double* out, int32_t* nrec);

/* Node relaxation with the stencil of the mesh generator's SOR
 * (setup_mesh_region, setup_mesh.F90:795-877; fixed weights :620-660).
 * type is the generator's region type: BL_RELAX_LIN1 averages the two
 * l-neighbours, BL_RELAX_LIN2 the two k-neighbours, BL_RELAX_EQUI uses the
 * equipotential (Winslow) weights of the node's 8 neighbours
 * (setup_mesh.F90:804-847).  A node moves by om*(stencil - x) when either
 * coordinate's residual exceeds tol (setup_mesh.F90:858-862).  Sweeps are
 * Jacobi sweeps: every node reads the previous sweep's positions (the
 * generator sweeps Gauss-Seidel in (l,k) order).  Nodes relaxed: interior
 * (indtype > 0) with four elements around them.  The logical frame of node
 * n is that of the lowest-numbered element with n at corner 1 (else the
 * lowest-numbered element around n): k toward that element's corner 2, l
 * toward its corner 4, as in the generator's meshes (setup_mesh_transfer). */
enum { BL_RELAX_OFF = 0, BL_RELAX_LIN1 = 1, BL_RELAX_LIN2 = 2, BL_RELAX_EQUI = 3 };
/* The mesh-motion option of the ALE step for zexist and not zeul, in the
 * empty "Other options" of ale_kn_getmeshstatus and ale_kn_getmeshvelocity
 * (ale_kn_getmeshstatus.f90:41-43, ale_kn_getmeshvelocity.f90:42-44): every
 * node is marked moving (indstatus = 2, as Eulerian), and the mesh velocity
 * is (x_R - x)/dt, x_R the positions after nsweep sweeps from the
 * Lagrangian positions x.  type BL_RELAX_OFF (the default) leaves the
 * reference's behaviour: no mesh motion option. */
int bl_set_relaxation(bl_handle_t* h, int32_t type, int32_t nsweep, double om, double tol);
/* nsweep sweeps on ndx, ndy in place (mesh set-up); the geometry is not
 * recomputed. */
int bl_relax_mesh(bl_handle_t* h, int32_t type, int32_t nsweep, double om, double tol);

/* ---- ensembles: many independent problems stepped together ----------
 * One instance: an nx x ny structured mesh on [0,length] x [0,height] with
 * reflecting walls (SLIPY bottom and top, SLIPX left and right), one
 * ideal-gas material of ratio gamma and two regions of anti-hourglass
 * coefficient kappa.  kind 0 (Sod): density and energy rho_l, e_l in the
 * elements whose corner mean has x < length/2 (region 1), rho_r, e_r
 * elsewhere (region 2).  kind 1 (Sedov): density rho_l everywhere, energy
 * e_l/(rho_l V) in the element at the origin (e_l is the deposited energy)
 * and 0 elsewhere, region 1.  Every instance needs nx*ny >= 256. */
typedef struct bl_instance {
  int32_t nx, ny, kind, pad;
  double length, height, gamma, kappa;
  double rho_l, e_l, rho_r, e_r;
} bl_instance_t;
/* A handle holding the disjoint union of the instances' meshes, generated
 * on the device with their initial state (setup_mesh_region's LIN1 block,
 * inf_init's geometry, mass and EOS set-up).  Instance k owns elements
 * eoff[k] .. eoff[k+1]-1 and nodes noff[k] .. noff[k+1]-1 of the
 * registry; instance k has material k+1 and regions 2k+1, 2k+2.  cfg is
 * shared by every instance (zexist must be 0). */
int bl_ensemble_create(const bl_instance_t* inst, int32_t ninst, const bl_config_t* cfg,
                       bl_handle_t** out);
/* nsteps steps of every instance, one captured hipGraph per step, each
 * instance with its own dt (time_dr_calc/end on its own CFL and DIV
 * controls).  The histories hold nsteps x ninst entries, step-major; idt
 * is the instance-local 1-based element.  A failure stops every instance;
 * err_idx names the element (negative volume) or instance (dt). */
int bl_ensemble_run(bl_handle_t* h, int32_t nsteps, double* dt_hist, int32_t* idt_hist,
                    int32_t* ctl_hist, int32_t* steps_taken, bl_error_t* err);
/* the offset tables (ninst + 1 entries each) */
int bl_ensemble_offsets(bl_handle_t* h, int32_t* eoff, int32_t* noff);
/* each instance's time, last dt and step count */
int bl_ensemble_times(bl_handle_t* h, double* time, double* dt, int32_t* nstep);

/* ---- learned node relaxation ------------------------------------------
 * No counterpart in the reference: a multilayer perceptron 26-64-64-2
 * (ReLU) that proposes each relaxed node's ALE displacement from its
 * stencil, run on the matrix cores with BF16 weights and activations and
 * FP32 accumulation, trained on (stencil, displacement) pairs the plain
 * relaxation (bl_set_relaxation) produces.
 *
 * Features of a relaxed node n with stencil nodes nb[0..7] (the
 * relaxation's table order) and h^2 the mean area of its four elements:
 * [0..15] (x_nb - x_n, y_nb - y_n)/h; [16..19] the element areas / h^2 and
 * [20..23] their corner Jacobians at n / h^2, elements counter-clockwise
 * from the one between nb[4] and nb[2]; [24..25] the displacement / h of
 * one sweep of the relaxation set by bl_set_relaxation.  Target: the
 * displacement / h after that relaxation's nsweep sweeps.  The network
 * learns the target's difference from [24..25]: a network that outputs zero
 * moves the mesh by one plain sweep.
 *
 * Parameters (FP32 master copy, 6018 floats): W1 (64,26) row-major, b1
 * (64), W2 (64,64), b2 (64), W3 (2,64), b3 (2); N(x) = W3 r(W2 r(W1 x +
 * b1) + b2) + b3.  With the standardisation (56 floats: input means mu[26],
 * inverse input standard deviations is[26], output means mt[2], output
 * standard deviations st[2]) the proposed displacement / h is
 * f[24..25] + mt + st * N((f - mu) * is), the standardised input rounded to
 * BF16.  The network's own device stream orders its calls. */
typedef struct bl_mlp bl_mlp_t;
enum { BL_MLP_NPARAM = 6018, BL_MLP_NIN = 26, BL_MLP_NOUT = 2, BL_MLP_NNORM = 56 };
/* He-uniform weights from seed, zero biases; room for capacity pairs */
int bl_mlp_create(uint64_t seed, int32_t capacity, bl_mlp_t** out);
int bl_mlp_destroy(bl_mlp_t* m);
/* copy the master parameters out (get) and/or in (set), either may be NULL */
int bl_mlp_weights(bl_mlp_t* m, float* get, const float* set);
/* the standardisation from the stored pairs' per-column mean and standard
 * deviation, of the features and of t - x[24..25] (a column whose deviation
 * is ~0 keeps scale 1); bl_mlp_train
 * calls it on its first use when no standardisation has been set */
int bl_mlp_normalize(bl_mlp_t* m);
/* copy the standardisation out (get) and/or in (set), either may be NULL */
int bl_mlp_norm(bl_mlp_t* m, float* get, const float* set);
/* y (n x 2) = the proposed displacement / h for features x (n x 26), as
 * the matrix cores evaluate the network */
int bl_mlp_forward(bl_mlp_t* m, const float* x, int32_t n, float* y);
/* append n pairs (x n x 26, t n x 2) from the host */
int bl_mlp_add_pairs(bl_mlp_t* m, const float* x, const float* t, int32_t n);
/* the pair count, and the pairs when x, t are not NULL */
int bl_mlp_pairs(bl_mlp_t* m, int32_t* n, float* x, float* t);
/* gradient of the loss mean((N(x_s) - t_s)^2) over pairs 0..n-1, x_s the
 * standardised inputs and t_s the standardised differences t - x[24..25],
 * with respect to the parameters
 * (the forward pass rounds weights, inputs and hidden activations to BF16
 * as the matrix cores see them), and the loss */
int bl_mlp_gradient(bl_mlp_t* m, int32_t n, float* grad, float* loss);
/* iters Adam steps (beta 0.9, 0.999, eps 1e-8, rate lr) on batches of
 * batch pairs drawn from the stored pairs by seed; loss_hist (iters) may be
 * NULL.  Deterministic for a given seed. */
int bl_mlp_train(bl_mlp_t* m, int32_t iters, int32_t batch, float lr, uint64_t seed,
                 float* loss_hist);
/* With m, every ALE step of h appends a pair for every stride-th relaxed
 * node (the plain relaxation's mesh motion of the step's Lagrangian mesh)
 * to m while they fit; NULL stops.  Needs bl_set_relaxation.  With
 * bl_relax_learned on as well, the network moves the mesh and the plain
 * motion is computed beside it for the pairs only (training on the states
 * the network leads to). */
int bl_relax_collect(bl_handle_t* h, bl_mlp_t* m, int32_t stride);
/* With m, the ALE mesh motion of h is the network's: x + h*y per relaxed
 * node, velocity h*y/dt.  The proposed mesh is checked element by element
 * (bl_set_learned_check); the four nodes of every failing element move by
 * one plain sweep instead (the guard) and the mesh is checked again; if it
 * still fails, the step uses the plain relaxation set by bl_set_relaxation
 * instead (the fallback).  NULL goes back to the plain relaxation.  m
 * must outlive its use by h.  The step counters of bl_relax_learned_stats
 * restart. */
int bl_relax_learned(bl_handle_t* h, bl_mlp_t* m);
/* The check of the proposed mesh: an element fails with a corner Jacobian
 * <= 0 or a longest/shortest side ratio > max_aspect.  relative 0: every
 * such element fails.  relative 1 (the default, with max_aspect 10): such
 * an element fails only when the proposal is worse (smaller corner
 * Jacobian, larger ratio) than the same element after one plain sweep from
 * the step's Lagrangian mesh, so that a poor element the Lagrangian motion
 * made, and which the plain relaxation keeps as well, does not force the
 * fallback; and an element the proposal shrinks below half its one-sweep
 * area fails too (a proposal that collapses elements would collapse dt). */
int bl_set_learned_check(bl_handle_t* h, double max_aspect, int32_t relative);
/* since the network was attached: ALE steps run with it, the steps that
 * used the guard, the steps that fell back, and the reasons seen (bit 0: a
 * corner Jacobian <= 0, bit 1: an aspect ratio > max_aspect, bit 2: an
 * element shrunk below half its one-sweep area); any output may be NULL */
int bl_relax_learned_stats(bl_handle_t* h, int32_t* steps, int32_t* guarded, int32_t* fallbacks,
                           int32_t* reasons);
/* the features of every node at the current positions (nnd x 26, zero
 * rows for nodes that are not relaxed) */
int bl_relax_features(bl_handle_t* h, float* feat);

/* Kernel timing with HIP events on the handle's stream: average duration
 * (ms) of the named kernel class over the launches recorded since the last
 * bl_timing_reset. */
int bl_timing_enable(bl_handle_t* h, int32_t on);
int bl_timing_reset(bl_handle_t* h);
int bl_timing_query(bl_handle_t* h, const char* name, double* avg_ms,
                    int64_t* launches);
/* comma-separated list of the kernel classes seen so far */
int bl_timing_names(bl_handle_t* h, char* buf, int32_t buflen);

/* Counters of the last bl_run_graph or bl_ensemble_run call (no
 * counterpart in the reference): moving = the elements whose nodes did not
 * all share one velocity, summed over the steps (the limiter pass's work);
 * zeroed = the elements at rest whose zero viscosity fields the step wrote,
 * summed over the steps.  An element at rest in two consecutive steps of a
 * call keeps the zeros of the first: neither written again nor read by the
 * force kernel.  Either output may be NULL. */
int bl_run_stats(bl_handle_t* h, int64_t* moving, int64_t* zeroed);

/* Launch-shape autotuner (no counterpart in the reference, whose loops have
 * no launch shape).  Times the limiter pass -- the viscosity and CFL of the
 * elements with moving nodes, the step's one worklist-driven kernel -- under
 * 15 shapes: 64, 128 or 256 threads per workgroup times 128..2048
 * workgroups (a single problem) or 1024..16384 elements per workgroup and
 * instance (an ensemble).  Each shape is timed twice, on the worklist of the
 * handle's current state and with every element listed (the worklist as the
 * flow spreads), each the median of reps launches after one warm-up; the
 * shape kept for every later step is the one whose worse ratio to the best
 * time of the two is smallest.  The state is left as it was (the viscosity fields the trials
 * write are restored) and no result depends on the choice: every reduction
 * of the pass is a first-index extremum over (value, index) pairs.  A handle
 * whose run has ended (time_end reached, or an error) is tuned as the next
 * run call would start it, and its device timestep is put back as it was.
 * out (may be NULL) receives the choice, the
 * number of shapes timed, and the chosen and the default shape's times. */
typedef struct {
  int32_t limiter_block; /* threads per workgroup */
  int32_t limiter_grid;  /* workgroups launched */
  int32_t ens_chunk;     /* ensemble: elements per workgroup and instance; 0 otherwise */
  int32_t trials;        /* shapes timed */
  double best_us;        /* median duration of the chosen shape, current worklist */
  double default_us;     /* the same for the shape in use before the call */
  double best_full_us;   /* median duration of the chosen shape, every element listed */
  double default_full_us; /* the same for the shape in use before the call */
} bl_tuning_t;
int bl_autotune(bl_handle_t* h, int32_t reps, bl_tuning_t* out);
/* Apply a launch shape without timing (one bl_autotune chose for the same
 * problem before, say): limiter_block (a multiple of 64 up to 256) and
 * limiter_grid (>= 1, a single problem) or ens_chunk (>= 64, an ensemble);
 * the other fields are ignored.  No result depends on it. */
int bl_set_tuning(bl_handle_t* h, const bl_tuning_t* t);
/* the shape in use: limiter_grid is the number of workgroups launched;
 * trials and the four times are zero */
int bl_get_tuning(bl_handle_t* h, bl_tuning_t* t);

/* The tile step (no counterpart in the reference, which has no kernels to
 * fuse).  A Lagrangian step of a problem with pure cells only and no
 * per-step diagnostics can run its predictor, node update and corrector
 * (the work of hydro_dr_getforce, getstate, getenergy, geteos and the
 * acceleration drivers) as one kernel per tile of about 180 elements and
 * the ring of elements around their nodes, keeping the corner forces and
 * node results in LDS: tiles come from recursive coordinate bisection of
 * the element centroids the first time bl_run_graph needs them after the
 * connectivity was built (bl_build_maps).  The mesh
 * must have no node that is the same corner (1..4) of two elements; a mesh
 * that does runs the three-kernel step.  on = 0 keeps the three-kernel step
 * for every problem.  No result depends on the choice: the sums over a
 * node's corners stay in the reference's order. */
int bl_set_tiling(bl_handle_t* h, int32_t on);
/* whether the next run uses the tile step (as far as known: the tiling is
 * built by the first run that needs it), the tiles, nboundary (always 0:
 * every element is completed by the tile that owns it; kept for callers of
 * the node tiling this replaced) and the tiles' element entries (each
 * element once in its own tile and once in the ring of every other tile
 * that holds one of its nodes); any output may be NULL */
int bl_get_tiling(bl_handle_t* h, int32_t* on, int32_t* ntile, int32_t* nboundary,
                  int32_t* nentries);

/* last error message of the library (thread-local), for diagnostics */
const char* bl_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BOOKLEAF_HIP_H */
