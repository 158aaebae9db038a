// bl_internal.h -- handle layout and launch helpers of the BookLeaf HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/bookleaf_hip.h"

namespace bl {

constexpr int kBlock = 256;  // 4 wave64 per workgroup
// workgroups of a grid-stride launch that usually exits at once (a gated
// fallback): one per CU
constexpr unsigned kGatedBlocks = 256;

void set_error(const std::string& msg);

#define BL_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::bl::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return BL_FAILURE;                                                          \
    }                                                                             \
  } while (0)

inline unsigned nblocks(int64_t n, int bs = kBlock) {
  return (unsigned)((n + bs - 1) / bs);
}

// dt reduction results written by the last workgroup of the getdt kernels
struct DtResult {
  double cfl_min;   // min over elements of rscratch11 (len^2/cs2)
  double div_max;   // max over elements of |div u|/V (TINY if none larger)
  int32_t cfl_idx;  // 1-based, first index attaining the min
  int32_t div_idx;  // 1-based, 0 if none exceeds TINY
  int32_t negvol;   // largest 1-based index with V < 0 (0 if none)
  int32_t negvol2;  // the same for the corrector geometry of the fused step
};

// device-resident timestep state of the fused step (bl_run_graph): the
// fields of timestep_t (dataAPI_types.f90:46-52) plus the batch control
struct DevTimestep {
  double time, dt, dts;
  double rdt;        // 1/dt, formed once where dt is selected (hydro_kn_getforce.f90:88)
  int32_t nstep, idtel, sdt, ale_on;
  int32_t done;      // time_end reached or an error: later steps do nothing
  int32_t skip;      // this step does nothing (set by the step prologue)
  int32_t err;       // 0 none, 1 negative volume, 2 negative CFL, 3 dt < dt_min
  int32_t err_step;  // nstep of the failing step
  int32_t err_idx;   // -iel of the last negative-volume element
  int32_t hist_n;    // steps recorded in the history buffer
  int32_t flips;     // steps whose node update wrote the other node set
  int32_t hist_base;  // ensembles: instance step count at the batch's start
  // ensembles: the smallest 4*instance + code over the instances whose dt
  // selection failed (code 2 negative CFL, 3 dt < dt_min), kNoPending when
  // none; written only by atomicMin in the dt selection, turned into
  // done/skip by the step's last kernel and into err by the host
  int32_t epend;
  // the tile step: the nstep of the batch's last step, and whether the step
  // running is the batch's last (it leaves its corner forces in
  // rscratch23/24, which its kernels otherwise keep in LDS)
  int32_t last_step;
  int32_t last;
  // a step of a pure-cell Lagrangian batch whose end (the negative-volume
  // outcome and the node-set flip) the next step's F1b, or the host after
  // the batch, still has to record
  int32_t open;
  // the node set the batch started from (the final one is set0 + flips)
  int32_t set0;
};
constexpr int32_t kNoPending = 0x7fffffff;

// per-launch kernel timing with HIP events on the handle's stream
struct TimingClass {
  std::string name;
  double total_ms = 0.0;
  int64_t launches = 0;
};

struct Handle {
  int device = 0;
  hipStream_t stream = nullptr;
  bl_sizes_t sz{};
  bl_config_t cfg{};
  // per-region / per-material tables (host + device)
  int32_t nreg = 0, nmat = 0;
  std::vector<double> kappareg, pmeritreg;
  std::vector<int32_t> zdtnotreg, zmidlength, eos_itype;
  std::vector<double> eos_param;
  double* d_kappareg = nullptr;
  double* d_pmeritreg = nullptr;
  int32_t* d_zdtnotreg = nullptr;
  int32_t* d_zmidlength = nullptr;
  int32_t* d_eos_itype = nullptr;
  double* d_eos_param = nullptr;
  bool ideal_single = false;  // every material 1..nmat ideal gas: fast EOS path
  // dataAPI field mirror
  void* f[BL_NFIELD + 1] = {};
  int32_t kind[BL_NFIELD + 1] = {};
  int64_t count[BL_NFIELD + 1] = {};
  // node -> corner CSR maps (corner id = 4*(iel-1) + (j-1))
  int32_t* d_hyd_ptr = nullptr;  // order (corner j, element in ielsort1 order)
  int32_t* d_hyd_idx = nullptr;
  int32_t* d_ale_ptr = nullptr;  // order (element in ielsort2 order, corner)
  int32_t* d_ale_idx = nullptr;
  int32_t* d_el_ptr = nullptr;   // order (element 1..nel, corner)
  int32_t* d_el_idx = nullptr;
  // limiter stencil of a consistently oriented conforming quad mesh: per
  // element, the far node of each of the 8 neighbour edges the Christensen
  // limiter reads (null when some element breaks that pattern)
  int32_t* d_outer = nullptr;
  bool maps_built = false;
  // reductions
  double* d_part = nullptr;  // 4 doubles per workgroup
  int32_t* d_parti = nullptr;
  uint32_t* d_ticket = nullptr;  // 4 counters
  DtResult* d_dtres = nullptr;
  DtResult* h_dtres = nullptr;  // pinned
  int64_t npart = 0;
  // fused-step state
  DevTimestep* d_ts = nullptr;
  DevTimestep* h_ts = nullptr;  // pinned
  double* d_hist = nullptr;     // dt history (3 values per step)
  // what a run call's end publishes for the host (pinned, coherent host
  // memory written by k_publish): the timestep and hist_cap history records
  DevTimestep* h_pub_ts = nullptr;
  double* h_pub_hist = nullptr;
  int64_t hist_cap = 0;
  double* d_tmp[8] = {};        // fused-step temporaries (nel or nnd sized)
  // node ping-pong of the fused step: set[0] is the registry's own ndx..ndv,
  // set[1] the alternate; cur_set names the one h->f[ndx..ndv] points at
  double* nodeset[2][4] = {};
  int cur_set = 0;
  // corrector corner terms of the fused step, corner-major (4*nel each):
  // force x, force y, half-step corner weight, nodal-mass term
  double* cterm[4] = {};
  // an ensemble's worklist (nel): elements whose nodes do not share one
  // velocity, one segment per instance
  int32_t* d_work = nullptr;
  uint64_t* d_rest = nullptr;             // F1's at-rest mask, 4 words per 256 elements
  unsigned long long* d_vstats = nullptr;  // counters of the last run call
  // remap temporaries of the fused step (zexist): 4*nel each for the face
  // and corner arrays, nel each for the element scalars
  double* ale_tmp[11] = {};
  // step graphs, each captured when a run call first needs it: [8 * fin +
  // 2 * si + p] runs 2^si steps (si = 0..3) from node set p, then with fin
  // the batch's end (k_batch_end, k_materialize, k_publish).  Ensembles use
  // [p], one step from node set p.
  hipGraphExec_t graph_exec[16] = {};
  // ensembles (bl_ensemble_create): instances stepped together
  int32_t ens_n = 0;
  std::vector<int32_t> ens_eoff, ens_noff;  // host copies, ens_n + 1 each
  int32_t* d_eoff = nullptr;
  int32_t* d_lboff = nullptr;      // F1L workgroup offsets per instance
  int32_t ens_lblocks = 0;
  uint16_t* d_einst = nullptr;
  uint16_t* d_ninst = nullptr;
  void* d_inst = nullptr;          // InstTs[ens_n]
  double* d_inst_dt = nullptr;
  uint32_t* d_cnt = nullptr;
  double* d_epart = nullptr;
  int32_t* d_eparti = nullptr;
  double* d_ehist = nullptr;
  int64_t ehist_cap = 0;
  // node relaxation as the ALE mesh motion (bl_set_relaxation)
  int32_t relax_type = 0, relax_nsweep = 1;
  double relax_om = 1.0, relax_tol = 0.0;
  int32_t* d_ring = nullptr;      // (8, nnd) stencil nodes, -1: not relaxed
  double* relax_tmp[4] = {};      // sweep ping-pong, nnd each
  // learned relaxation (bl_relax_learned / bl_relax_collect)
  struct bl_mlp* mlp_infer = nullptr;
  struct bl_mlp* mlp_collect = nullptr;
  int32_t collect_stride = 1;     // every stride-th relaxed node per step
  double* d_cw[2] = {};           // plain motion collected beside the learned one
  double learned_max_aspect = 10.0;  // the check of the proposed mesh
  int32_t learned_relative = 1;
  int32_t* d_rlist = nullptr;     // relaxed nodes, ascending
  int32_t nrelaxed = 0;
  int32_t* d_zlist = nullptr;     // nodes that are not relaxed, ascending
  int32_t nzero = 0;
  int32_t* d_lflag = nullptr;     // learned-motion flags and counters
  int32_t* d_lmark = nullptr;     // nnd: node marked by the check for the guard
  int32_t* d_nbfc = nullptr;      // (4, nel) ielel/ielfc packed for the fused remap
  // launch shape of the limiter pass, set by bl_autotune: threads per
  // workgroup, workgroups (single problem), elements per workgroup and
  // instance (ensembles)
  int32_t tune_lim_block = 256, tune_lim_grid = 512, tune_ens_chunk = 4096;
  int32_t ens_lblocks_cap = 0;    // limiter partial slots allocated
  bool nbfc_stale = true;         // rebuild it before the next capture
  // the tile step of a pure-cell Lagrangian problem (bl_tiles.hip): the
  // elements cut into tiles, each with the nodes of its own elements (at
  // most kTileNodes) and the elements around those nodes (its own, then the
  // ring; at most kTileEls)
  int32_t tile_state = 0;         // 0 not built, 1 built, -1 the mesh has no tiling
  bool tiles_on = true;           // bl_set_tiling
  int32_t ntile = 0, nentries = 0;
  int32_t* d_tile_eptr = nullptr;  // ntile + 1
  int32_t* d_tile_nptr = nullptr;  // ntile + 1
  int4* d_tile_bounds = nullptr;   // per tile (eptr[t], eptr[t+1], nptr[t], nptr[t+1])
  int32_t* d_tile_nodes = nullptr; // the tile's nodes, 0-based | (it writes the node) << 31
  uint8_t* d_tile_ncm = nullptr;   // beside d_tile_nodes: bit jj, the node is some element's corner jj
  int2* d_tile_el = nullptr;       // (element | flags << 30, four local node slots)
  int4* d_tile_nd = nullptr;       // the entries' four nodes, 0-based
  int32_t* d_ens_toff = nullptr;   // ensembles: ens_n + 1 tile offsets per instance
  uint8_t* d_rest8 = nullptr;      // 2 x nel at-rest bytes of the tile step
  // the element state of the tile step in two sets (eldensity, elenergy,
  // elpressure, elcs2); el_cur names the one the registry points at
  double* elset[2][4] = {};
  int el_cur = 0;
  float* d_ld1 = nullptr;         // (nnd, 2): one-sweep displacement of the node
  // per-step diagnostics of the fused step: region table + Sod L1 records
  bool diag_on = false;
  double diag_x0 = 0.0, diag_scale = 1.0;
  int32_t diag_max = 0;
  double* d_diag_hist = nullptr;     // diag_max records of 2 + 9*nreg + 2
  int32_t* d_diag_count = nullptr;   // records written
  double* d_diag_part = nullptr;     // per-workgroup partials
  size_t diag_part_n = 0;
  int graph_stop_at_end = -1;
  // The next bl_run_graph may continue from the last one's end: it keeps
  // the at-rest masks and the F1 terms the last step left for the next
  // step instead of clearing the masks and running F1 again.  Set by a run
  // call that ran every step of a folded pure-cell Lagrangian batch; cleared
  // by every entry point that can write device state (bl_touch).  A device
  // pointer handed out (bl_field_device_ptr) allows writes this library
  // does not see, so it ends resuming for the handle's lifetime.
  bool resumable = false;
  bool ptr_out = false;
  // timing
  bool timing = false;
  std::vector<TimingClass> tclass;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_class;  // class of each event pair in the pool
  int ev_used = 0;

  double* R(int id) const { return (double*)f[id]; }
  int32_t* I(int id) const { return (int32_t*)f[id]; }
  uint8_t* Z(int id) const { return (uint8_t*)f[id]; }
};

// timing helpers (bl_api.hip)
int timing_begin(Handle* h, const char* cls);
void timing_end(Handle* h, int slot);
// lazily allocated nsz-sized temporaries (bl_ale.hip)
double* tmp_buffer(Handle* h, int k);
// ALE step launches; with d_ts the kernels read dt, the step gate and
// ale_on from the device timestep (bl_ale.hip)
int ale_step_device(Handle* h, const DevTimestep* d_ts);
// the relaxation's ALE mesh velocity (x_R - x)/dt into (wu, wv), gated by
// ts when given; the stencil table rebuilt from the connectivity
int relax_velocity(Handle* h, const double* x, const double* y, double* wu, double* wv,
                   double dt, const DevTimestep* ts, const int32_t* only_if = nullptr);
// the learned relaxation's mesh velocity (bl_relax_learned), with the plain
// relaxation as the fallback; and the training pairs of the plain relaxation
// from x (Lagrangian) and its mesh velocity (bl_relax_collect)
int learned_velocity(Handle* h, const double* x, const double* y, double* wu, double* wv,
                     double dt, const DevTimestep* ts);
int collect_pairs(Handle* h, const double* x, const double* y, const double* wu,
                  const double* wv, double dt, const DevTimestep* ts);
// the ALE mesh velocity of the relaxation in use (learned or plain), and
// the training pairs when collecting
int mesh_motion(Handle* h, const double* x, const double* y, double* wu, double* wv, double dt,
                const DevTimestep* ts);
int relax_rebuild(Handle* h);
// the per-step diagnostics record from the node set a step wrote (and the
// element state it wrote: the registry's unless given)
int launch_diag(Handle* h, const double* ndx, const double* ndy, const double* ndu,
                const double* ndv, const double* rho = nullptr, const double* en = nullptr,
                const double* pre = nullptr);
// ensembles: the limiter pass's workgroups per instance for a chunk of
// elements each (bl_ensemble.hip)
int ens_set_chunk(Handle* h, int chunk);
// drop the captured step graphs (config or maps changed)
void invalidate_graph(Handle* h);

}  // namespace bl

struct bl_handle : public bl::Handle {};

// an entry point that may write the handle's device state: the next run
// call starts afresh
inline void bl_touch(bl_handle* h) {
  if (h) h->resumable = false;
}

// launch a kernel on the handle's stream, timed by class name when enabled
#define BL_LAUNCH(h, cls, kern, grid, block, ...)                                  \
  do {                                                                             \
    int slot_ = (h)->timing ? ::bl::timing_begin((h), (cls)) : -1;                 \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, (h)->stream, __VA_ARGS__); \
    if (slot_ >= 0) ::bl::timing_end((h), slot_);                                  \
    hipError_t le_ = hipGetLastError();                                            \
    if (le_ != hipSuccess) {                                                       \
      ::bl::set_error(std::string("launch ") + #kern + ": " + hipGetErrorString(le_)); \
      return BL_FAILURE;                                                           \
    }                                                                              \
  } while (0)
