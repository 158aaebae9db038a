// bl_fused.hip -- the device-resident whole step (bl_run_graph).
//
// One step (inf_solver_getdt + inf_solver_lagstep [+ inf_solver_alestep])
// as a few grid kernels and one-workgroup kernels, with dt, time and the
// error state kept on the device, so that a step is captured as a hipGraph
// and replayed with no host synchronisation:
//
//   F1 getdt   : every element: corner gather, the DIV term, whether its four
//                nodes share one velocity (then its zero viscosity and CFL
//                term here), the at-rest mask, per-workgroup partials
//   F1L limiter: the elements in motion (from the mask): the limiter and
//                edge viscosity (neighbour edges from node data), cnviscx/y,
//                elvisc, CFL term; folds F1's partials into its own
//   F1b select : one workgroup: the step before's end, first-index min/max
//                over the partials and time_dr_calc / time_dr_end -> dt,
//                1/dt, time, nstep on the device
//   F2 force   : predictor force, half-step geometry, energy and EOS in
//                registers, then the corrector's corner forces P+Q(+SP)(+HG),
//                half-step corner weights and nodal-mass terms, corner-major
//                (X[j*nel+e]) so that F3 reads each of them once, coalesced
//   F3 accel   : node gather through the (corner, element) CSR map in the
//                reference's summation order -> acceleration, BCs, u^{n+1},
//                ubar, x^{n+1} into the other node set (ping-pong)
//   F4 state   : corrector geometry, density, energy with ubar, EOS; in a
//                Lagrangian step of pure cells also the next step's F1 terms
//                (F1 then runs once per run call)
//   F5 end     : errors and the node-set flip, when no later kernel of the
//                step or the next F1b can record them; F2m, F4m: the mixed
//                cells' component terms, in F2's and F5's extra workgroups
//   tile step  : in a Lagrangian step of pure cells, F2, F3 and F4 in one
//                kernel per tile of elements and the ring around them
//                (k_lag_tile, bl_tiles.hip)
//   R1..R5     : the remap (ale_dr_*), after the Lagrangian step when ALE is
//                active
//
// Every FP64 expression is the same function of the same inputs as in the
// per-driver path (bl_math.h), so the results are bit-identical to the
// reference drivers; the intermediate half-step pressure/energy never reach
// HBM.  Fields that hold only phase scratch in the reference (rscratch*) are
// not written except ubar (rscratch14/15, which the ALE step reads) and the
// corrector corner forces (rscratch23/24): the tile step's last step of a
// batch writes them, otherwise k_materialize does at the end of a batch from
// the last step's corner-major copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <type_traits>
#include <vector>

#include "bl_fused.h"
#include "bl_ale.h"
#include "bl_internal.h"
#include "bl_kernels.h"
#include "bl_math.h"
#include "bl_tiles.h"

namespace bl {

// whether this step does nothing: the batch is done (an error, or time_end
// reached), or stop_at_time_end and time >= time_end (inf_control_solver's
// loop test).  F1 and F1b decide it from the timestep as the step begins;
// F1b then records it in ts->skip for the kernels that follow.
__device__ __forceinline__ bool step_skipped(const FusedArgs& A, const DevTimestep* ts) {
  return ts->done || (A.stop_at_time_end && ts->time >= A.time_end);
}

// the limiter pass's at-rest flags per thread and scan pass (a byte word)
constexpr int kScanPer = 8;
// F1b's workgroup: four partials of each control per thread at the usual
// 1024 limiter workgroups, all loaded at once (512 threads were 1-1.5 us a
// step faster than 1024, which take longer to start, and 256 another
// 0.9 us faster than 512)
constexpr int kF1bThreads = 256;

// a wave's count of elements whose zero viscosity it wrote, to its
// workgroup's counter slot
__device__ __forceinline__ void count_zeroed(const FusedArgs& A, uint64_t wrote) {
  atomicAdd(A.vstats + kStatStride * (1 + (blockIdx.x & (kStatSlots - 1))),
            (unsigned long long)__popcll(wrote));
}

__device__ __forceinline__ void load4(const double* __restrict__ a, const int nd[4],
                                      double v[4]) {
  v[0] = a[nd[0]]; v[1] = a[nd[1]]; v[2] = a[nd[2]]; v[3] = a[nd[3]];
}
__device__ __forceinline__ void load_nd(const int32_t* __restrict__ ielnd, int e, int nd[4]) {
  int4 q = reinterpret_cast<const int4*>(ielnd)[e];
  nd[0] = q.x - 1; nd[1] = q.y - 1; nd[2] = q.z - 1; nd[3] = q.w - 1;
}
__device__ __forceinline__ void st4(double* __restrict__ a, int e, const double v[4]) {
  reinterpret_cast<double4*>(a)[e] = make_double4(v[0], v[1], v[2], v[3]);
}
// a streaming store, for results the step's kernels do not read again (the
// batch's end writes them once): the cache keeps the step's working set
typedef double nt_double4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st4_nt(double* __restrict__ a, int e, const double v[4]) {
  const nt_double4 q = {v[0], v[1], v[2], v[3]};
  __builtin_nontemporal_store(q, reinterpret_cast<nt_double4*>(a) + e);
}
__device__ __forceinline__ void st_nt(double* __restrict__ a, int e, double v) {
  __builtin_nontemporal_store(v, a + e);
}
__device__ __forceinline__ void ld4(const double* __restrict__ a, int e, double v[4]) {
  double4 q = reinterpret_cast<const double4*>(a)[e];
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}

// a node's (u, v, x, y)
struct NodeUVXY {
  double u, v, x, y;
};
__device__ __forceinline__ NodeUVXY node_uvxy(const FusedNodes& N, int a) {
  return NodeUVXY{N.u[a], N.v[a], N.x[a], N.y[a]};
}
// differences (b - a) of node data along the edge a -> b
__device__ __forceinline__ void edge_diff(const NodeUVXY& a, const NodeUVXY& b, double& du,
                                          double& dv, double& dx, double& dy) {
  du = b.u - a.u;
  dv = b.v - a.v;
  dx = b.x - a.x;
  dy = b.y - a.y;
}

// ---------------------------------------------------------------- F1
// The limiter (hydro_kn_limiteviscosity) and edge viscosity
// (hydro_kn_geteviscosity) of element e: corner forces (qx, qy) and elvisc
// (ev, 0.0 on entry).  The four neighbour edges of each limiter direction
// come from the outer-node table: one end is a node of e, already in
// registers, the other is the neighbour's far node (a mesh without that
// structure runs host-driven).
__device__ __forceinline__ void f1_viscosity(const FusedArgs& A, const FusedNodes& N, int e,
                                             const int nd[4], const double x[4],
                                             const double y[4], const double u[4],
                                             const double v[4], double cs2e, double rhoe,
                                             double& ev, double qx[4], double qy[4]) {
  const int4 nb4 = reinterpret_cast<const int4*>(A.ielel)[e];
  const int nb[4] = {nb4.x - 1, nb4.y - 1, nb4.z - 1, nb4.w - 1};
  const int4 o0 = reinterpret_cast<const int4*>(A.outer)[2 * e];
  const int4 o1 = reinterpret_cast<const int4*>(A.outer)[2 * e + 1];
  const int on[8] = {o0.x - 1, o0.y - 1, o0.z - 1, o0.w - 1,
                     o1.x - 1, o1.y - 1, o1.z - 1, o1.w - 1};
  double ou[4], ov[4], ox[4], oy[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    int g = (f + 1) & 3;
    ou[f] = u[g] - u[f]; ov[f] = v[g] - v[f]; ox[f] = x[g] - x[f]; oy[f] = y[g] - y[f];
  }
  // the element's own nodes come from registers (the same values its
  // corners loaded); only the outer nodes are gathered, the four of a
  // limiter direction together
  NodeUVXY own[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) own[k] = NodeUVXY{u[k], v[k], x[k], y[k]};
  double ed[4];
#pragma unroll
  for (int iside = 0; iside < 2; ++iside) {
    const int is1 = (iside + 3) & 3, is2 = iside + 1;
    const int in1 = nb[iside], in2 = nb[iside + 2];
    double s1, s2, s3, s4;
    double au, av, ax, ay, bu, bv, bx, by;
    const int* o = on + 4 * iside;
    NodeUVXY out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = node_uvxy(N, o[k]);
    // in1 edge f1+1, in2 edge f2+3
    edge_diff(own[iside], out[0], au, av, ax, ay);
    edge_diff(out[2], own[(iside + 3) & 3], bu, bv, bx, by);
    limiter_ratios(ou[is1], ov[is1], ox[is1], oy[is1], au, av, ax, ay, bu, bv, bx, by,
                   A.zerocut, s1, s2);
    // in1 edge f1+3, in2 edge f2+1
    edge_diff(out[1], own[iside + 1], au, av, ax, ay);
    edge_diff(own[iside + 2], out[3], bu, bv, bx, by);
    limiter_ratios(ou[is2], ov[is2], ox[is2], oy[is2], au, av, ax, ay, bu, bv, bx, by,
                   A.zerocut, s3, s4);
    if (in1 == e) {
      int ic1 = nd[iside], ic2 = nd[(iside + 1) & 3];
      if (A.indtype[ic1] < 0 && A.indtype[ic2] < 0 && in2 != e) { s1 = 1.0; s3 = 1.0; }
      else { s1 = 0.0; s3 = 0.0; }
    }
    if (in2 == e) {
      int ic1 = nd[iside + 2], ic2 = nd[(iside + 3) & 3];
      if (A.indtype[ic1] < 0 && A.indtype[ic2] < 0 && in1 != e) { s2 = 1.0; s4 = 1.0; }
      else { s2 = 0.0; s4 = 0.0; }
    }
    int ins = nb[is1];
    ed[is1] = limiter_edge(s1, s2, ou[is1], ov[is1], cs2e, A.cs2[ins], rhoe, A.rho[ins],
                           A.cvisc1, A.cvisc2, A.zerocut, ev);
    ins = nb[is2];
    ed[is2] = limiter_edge(s3, s4, ou[is2], ov[is2], cs2e, A.cs2[ins], rhoe, A.rho[ins],
                           A.cvisc1, A.cvisc2, A.zerocut, ev);
  }
  geteviscosity(x, y, u, v, ed, ed, A.zerocut, qx, qy);
}

// the region switches of e's CFL term (zdtnotreg, zmidlength)
struct CflIn {
  int dtnot, mid;
};
// (of region ireg, 0-based)
// (loads either way, from the element's material when there are no region
// tables, so that no branch joins behind them)
__device__ __forceinline__ CflIn cfl_of_region(const FusedArgs& A, int ireg) {
  const int k = A.reg_tables ? ireg : 0;
  const int dtnot = (A.reg_tables ? A.zdtnotreg : A.ielmat)[k];
  const int mid = (A.reg_tables ? A.zmidlength : A.ielmat)[k];
  return CflIn{A.reg_tables ? dtnot : 0, A.reg_tables ? mid : 0};
}
__device__ __forceinline__ CflIn cfl_in(const FusedArgs& A, int e) {
  return cfl_of_region(A, A.reg_tables ? A.ielreg[e] - 1 : 0);
}

// the CFL term of e (hydro_kn_getdtcfl with cs2 + elvisc, hydro_dr_getcs2;
// a mixed cell takes the frmass average of its components instead)
__device__ __forceinline__ double f1_cfl(const FusedArgs& A, int e, const double x[4],
                                         const double y[4], double cs2e, double ev,
                                         CflIn r) {
  double c2 = cs2e + ev;
  if (A.mx.ncp > 0) {
    const int im = A.ielmat[e];
    if (im < 0) c2 = mix_cs2(A.mx, -im - 1);
  }
  return r.dtnot ? DBL_MAX : minlength(x, y, A.zcut, r.mid != 0) / c2;
}
__device__ __forceinline__ double f1_cfl(const FusedArgs& A, int e, const double x[4],
                                         const double y[4], double cs2e, double ev) {
  return f1_cfl(A, e, x, y, cs2e, ev, cfl_in(A, e));
}

// the wave's word of the at-rest mask, and the count of elements whose
// zeros F1 wrote (one atomic per wave that wrote any)
__device__ __forceinline__ void rest_update(const FusedArgs& A, int e, int lane, uint64_t prev,
                                            bool at_rest, bool zeroed) {
  const uint64_t now = __ballot(at_rest);
  const uint64_t wrote = __ballot(zeroed);
  if (lane == 0) {
    if (now != prev) A.rest[e >> 6] = now;
    if (wrote) count_zeroed(A, wrote);
  }
}

// F1: every element.  Four nodes with one velocity make every own-edge
// velocity difference zero, so geteviscosity zeroes every edge (|du| <=
// zerocut) whatever the limiter gives, and each edge's elvisc term is
// w3 * 0 = +0 (finite states): such an element gets its zero viscosity and
// CFL term here, and its bit in the at-rest mask.  Every other element is
// left to F1L (which finds it from the mask) for its viscosity and CFL term;
// its CFL slot here is DBL_MAX.  The DIV term of every element is computed
// here.  The CFL and DIV selections are first-index minima/maxima over
// (value, index) pairs, so the split and the order do not change the result.

// F1's terms of element e from the state a step starts from (nodes x, y, u,
// v; g the geometry at x; cs2e): whether its nodes move, its CFL term when
// they do not (DBL_MAX otherwise: F1L computes it), its DIV term
// (hydro_kn_getdtdiv)
__device__ __forceinline__ void f1_terms(const FusedArgs& A, int e, const double x[4],
                                         const double y[4], const double u[4], const double v[4],
                                         const Iso& g, double cs2e, CflIn r, bool& moving,
                                         double& cflv, double& divv) {
  moving = !(u[0] == u[1] && u[0] == u[2] && u[0] == u[3] && v[0] == v[1] && v[0] == v[2] &&
             v[0] == v[3]);
  if (!moving) cflv = f1_cfl(A, e, x, y, cs2e, 0.0, r);
  divv = divterm(u, v, g.a1, g.a3, g.b1, g.b3, g.vol);
}
__device__ __forceinline__ void f1_terms(const FusedArgs& A, int e, const double x[4],
                                         const double y[4], const double u[4], const double v[4],
                                         const Iso& g, double cs2e, bool& moving, double& cflv,
                                         double& divv) {
  f1_terms(A, e, x, y, u, v, g, cs2e, cfl_in(A, e), moving, cflv, divv);
}

// the workgroup's DIV and CFL partials, at part[blockIdx.x] and
// part[2 * gridDim.x + blockIdx.x] (the elements in motion are the at-rest
// mask's zeros, which F1L reads); every thread of the workgroup takes part
__device__ __forceinline__ void f1_collect(const FusedArgs& A, int e, int lane, bool moving,
                                           double cflv, double divv) {
  const int nb = gridDim.x;
  block_argmax_first(divv, e + 1, A.part, A.parti);
  __syncthreads();
  block_argmin_first(cflv, e + 1, A.part + 2 * nb, A.parti + 2 * nb);
}

// F1 in a step without the fold into F4, and as the first F1 of a run call
// with it (A.fold: F2 then writes the zeros)
__global__ void __launch_bounds__(kBlock)
k_f1_getdt(FusedArgs A, FusedNodes N) {
  if (step_skipped(A, A.ts)) return;
  const int e = xcd_index();  // a wave holds 64 consecutive elements from a multiple of 64
  const int lane = threadIdx.x & 63;
  const uint64_t prev = A.rest[e >> 6];
  double cflv = DBL_MAX, divv = -1.0;
  bool moving = false, zeroed = false;
  if (e < A.nel) {
    int nd[4];
    load_nd(A.ielnd, e, nd);
    double x[4], y[4], u[4], v[4];
    load4(N.x, nd, x); load4(N.y, nd, y); load4(N.u, nd, u); load4(N.v, nd, v);
    Iso g;
    getiso(x, y, g);
    f1_terms(A, e, x, y, u, v, g, A.cs2[e], moving, cflv, divv);
    // at rest in the previous step of this run too: its zeros are in place
    if (!moving && !A.fold && !((prev >> lane) & 1ull)) {
      const double z[4] = {0.0, 0.0, 0.0, 0.0};
      st4(A.cnviscx, e, z);
      st4(A.cnviscy, e, z);
      A.elvisc[e] = 0.0;
      zeroed = true;
    }
  }
  rest_update(A, e, lane, prev, e < A.nel && !moving, zeroed);
  f1_collect(A, e, lane, moving, cflv, divv);
}

// every element on an ensemble's worklist (the autotuner's second trial):
// element i at place i, which lies in its instance's segment
__global__ void __launch_bounds__(kBlock) k_work_all(int32_t* work, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) work[i] = i;
}

// viscosity, elvisc and CFL term of an element in motion (a NaN CFL term
// becomes DBL_MAX: it never wins the strict '<' scan)
__device__ __forceinline__ double f1_moving(const FusedArgs& A, const FusedNodes& N, int e) {
  int nd[4];
  load_nd(A.ielnd, e, nd);
  double x[4], y[4], u[4], v[4];
  load4(N.x, nd, x); load4(N.y, nd, y); load4(N.u, nd, u); load4(N.v, nd, v);
  const double cs2e = A.cs2[e], rhoe = A.rho[e];
  double ev = 0.0, qx[4], qy[4];
  f1_viscosity(A, N, e, nd, x, y, u, v, cs2e, rhoe, ev, qx, qy);
  st4(A.cnviscx, e, qx);
  st4(A.cnviscy, e, qy);
  A.elvisc[e] = ev;
  const double c = f1_cfl(A, e, x, y, cs2e, ev);
  return (c == c) ? c : DBL_MAX;
}

// time_dr_calc / time_dr_end's selection (time_dr_advance.f90:38-132):
// strict '<' minimum over MAXIMUM, GROWTH, INITIAL (first step), CFL, DIV
__device__ __forceinline__ void select_dt(const FusedArgs& A, double dt_prev, int nstep,
                                          double cfl_min, int cfl_idx, double div_max,
                                          int div_idx, double& best, int& bidx, int& bsdt) {
  best = A.dt_max;
  bidx = -1;
  bsdt = BL_DT_MAXIMUM;
  double c = A.dt_g * dt_prev;
  if (c < best) { best = c; bidx = -1; bsdt = BL_DT_GROWTH; }
  if (nstep == 0 && A.dt_initial < best) { best = A.dt_initial; bidx = -1; bsdt = BL_DT_INITIAL; }
  c = A.cfl_sf * sqrt(cfl_min);
  if (c < best) { best = c; bidx = cfl_idx; bsdt = BL_DT_CFL; }
  c = A.div_sf / div_max;
  if (c < best) { best = c; bidx = div_idx; bsdt = BL_DT_DIV; }
}

// The end of a step that ran (k_f5_end's work, when no F5 follows F4): a
// negative volume of the predictor (F2: the nodes were not moved) or of the
// corrector (F4: they were) stops the batch at that step, as
// geometry_kn_checkvolume's error halts the reference; a step that moved
// the nodes wrote the other node set.  Host and device form of the same
// record (the host runs it for a batch's last step).
__host__ __device__ inline void step_end(DevTimestep* ts, const DtResult* res) {
  if (!ts->open) return;
  ts->open = 0;
  if (res->negvol) {
    ts->err = 1; ts->err_idx = -res->negvol; ts->done = 1; ts->skip = 1;
    ts->err_step = ts->nstep;
    return;
  }
  ts->flips = ts->flips + 1;
  if (res->negvol2) {
    ts->err = 1; ts->err_idx = -res->negvol2; ts->done = 1; ts->skip = 1;
    ts->err_step = ts->nstep;
  }
}

// F1L: the elements in motion, found from the step's at-rest mask (bits
// in A.rest, or the tile step's bytes in A.rest8) over this workgroup's run
// of element blocks (its XCD's part of the mesh, as xcd_block deals them),
// compacted chunk by chunk in LDS so that every lane of a pass has an
// element, in element order.  Each workgroup also folds a slice of the nb1
// partials of F1 (or of the step kernels that did F1's work) of each
// control into its own, so that F1b reduces one (CFL, DIV) pair per F1L
// workgroup: CFL at part + 3 * nb1, DIV at part + nb1 (first-index extrema
// are associative, so the result is F1b's over all partials)
__global__ void __launch_bounds__(kBlock)
k_f1_limiter(FusedArgs A, FusedNodes N, int nb1) {
  if (step_skipped(A, A.ts)) return;
  // the step before found a negative volume that F1b has not recorded yet:
  // this step does nothing (F1b makes it so)
  if (A.end_in_f1b && (A.dtres->negvol || A.dtres->negvol2)) return;
  __shared__ int32_t list[kScanPer * kBlock];
  __shared__ uint32_t wave_n[kBlock / 64];
  double best = DBL_MAX, dv = -1.0;
  int bi = 0x7fffffff, di = 0x7fffffff;
  const int per = (nb1 + (int)gridDim.x - 1) / (int)gridDim.x;
  const int lo = (int)blockIdx.x * per, hi = min(nb1, lo + per);
  // this thread's first partial of each control, loaded now (from a clamped
  // index) and compared after the scan, so that the scan's first loads do
  // not wait for these
  const int k0 = lo + (int)threadIdx.x, kc = min(k0, nb1 - 1);
  const double fdv = A.part[kc], fcv = A.part[2 * nb1 + kc];
  const int fdi = A.parti[kc], fci = A.parti[2 * nb1 + kc];
  const int nbe = (A.nel + kBlock - 1) / kBlock;
  const int lb = xcd_block();
  const int e0 = (int)(((long long)lb * nbe) / gridDim.x) * kBlock;
  const int e1 = min(A.nel, (int)(((long long)(lb + 1) * nbe) / gridDim.x) * kBlock);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t listed = 0;
  // a pass covers kScanPer elements per thread (from a multiple of 8, so a
  // thread's flags are one aligned 8-byte word of bytes, or 8 bits of one
  // 64-bit word) and lists all the pass's moving elements before any
  // limiter work, so that a range with few of them costs one scan and one
  // limiter pass
  const int span = kScanPer * (int)blockDim.x;
  for (int c = e0; c < e1; c += span) {
    const int b = c + kScanPer * (int)threadIdx.x;  // this thread's first element
    uint32_t mv = 0;                                 // bit k: element b + k moves
    if (b < e1) {
      if (A.rest8) {
        if (b + kScanPer <= e1) {
          const uint64_t w = *reinterpret_cast<const uint64_t*>(A.rest8 + b);
#pragma unroll
          for (int k = 0; k < kScanPer; ++k) mv |= (((w >> (8 * k)) & 0xffull) == 0) ? 1u << k : 0u;
        } else {
          for (int k = 0; b + k < e1; ++k) mv |= A.rest8[b + k] == 0 ? 1u << k : 0u;
        }
      } else {
        const uint32_t rest = (uint32_t)(A.rest[b >> 6] >> (b & 63)) & 0xffu;
        mv = ~rest & 0xffu;
        if (b + kScanPer > e1) mv &= (1u << (e1 - b)) - 1u;
      }
    }
    // the block's exclusive prefix of the counts: within the wave by
    // shuffles, then over the waves
    const uint32_t cnt = (uint32_t)__popc(mv);
    uint32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    if (lane == 63) wave_n[wid] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (int w = 0; w < nw; ++w) {
      base += w < wid ? wave_n[w] : 0u;
      total += wave_n[w];
    }
    uint32_t at = base + incl - cnt;
    for (uint32_t m = mv; m; m &= m - 1) list[at++] = b + __ffs(m) - 1;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
      const int el = list[i];
      better_min(best, bi, f1_moving(A, N, el), el + 1);
    }
    listed += total;
    __syncthreads();  // the list is refilled
  }
  if (threadIdx.x == 0 && listed) atomicAdd(A.vstats, (unsigned long long)listed);
  if (k0 < hi) {
    better_max(dv, di, fdv, fdi);
    better_min(best, bi, fcv, fci);
  }
  for (int k = k0 + (int)blockDim.x; k < hi; k += (int)blockDim.x) {
    better_max(dv, di, A.part[k], A.parti[k]);
    better_min(best, bi, A.part[2 * nb1 + k], A.parti[2 * nb1 + k]);
  }
  // (one pass of shuffles and one barrier for both)
  if (!(best == best)) best = DBL_MAX;  // a NaN never wins the strict '<' scan
  block_reduce_div_cfl(dv, di, best, bi, A.part + nb1 + blockIdx.x, A.parti + nb1 + blockIdx.x,
                       A.part + 3 * nb1 + blockIdx.x, A.parti + 3 * nb1 + blockIdx.x);
}

// F1b: one workgroup -- combine the partials of F1 and F1L, then the
// hydro_dr_getdt candidates and time_dr_calc / time_dr_end
// (time_dr_advance.f90:38-132).  (A separate launch: handing this to F1L's
// last workgroup needs a device-scope fence in every workgroup, which costs
// far more than the launch with an L2 per XCD.)
__global__ void __launch_bounds__(kFinalBlock) k_f1b_select(FusedArgs A, int nb, int nbl) {
  // F1L's partials first: their loads do not wait for the gate below
  double va, vc;
  int xa, xc;
  final_dt_gather(nbl, A.part + nb, A.parti + nb, nbl, A.part + 3 * nb, A.parti + 3 * nb, va, xa,
                  vc, xc);
  DevTimestep* ts = A.ts;
  if (A.end_in_f1b) {
    if (threadIdx.x == 0) step_end(ts, A.dtres);
    __syncthreads();
  }
  const bool skip = step_skipped(A, ts);
  __syncthreads();  // every thread has read ts before thread 0 writes it
  if (threadIdx.x == 0) {
    if (skip) ts->done = 1;
    ts->skip = skip ? 1 : 0;
    A.dtres->negvol = 0;
    A.dtres->negvol2 = 0;
    if (A.end_in_f1b && !skip) ts->open = 1;
  }
  if (skip) return;
  final_dt_combine(va, xa, vc, xc, A.dtres);
  if (threadIdx.x != 0) return;
  const DtResult& r = *A.dtres;
  if (r.cfl_min < 0.0) {  // hydro_kn_getdt.f90:75-78
    ts->err = 2; ts->done = 1; ts->skip = 1; ts->err_step = ts->nstep + 1;
    return;
  }
  double best;
  int bidx, bsdt;
  select_dt(A, ts->dt, ts->nstep, r.cfl_min, r.cfl_idx, r.div_max, r.div_idx, best, bidx, bsdt);
  ts->nstep = ts->nstep + 1;
  ts->dt = best;
  ts->rdt = 1.0 / best;  // hydro_kn_getforce.f90:88 (w4 = 1/dt, once per call)
  ts->idtel = bidx;
  ts->sdt = bsdt;
  ts->time = ts->time + best;
  ts->dts = best;
  // the batch's last step (the one after it would be skipped)
  ts->last = (ts->nstep == ts->last_step) || (A.stop_at_time_end && ts->time >= A.time_end);
  ts->ale_on = A.zexist && (ts->time >= A.mintime) && (ts->time <= A.maxtime);
  if (A.hist && ts->hist_n < A.hist_cap) {
    int k = ts->hist_n++;
    A.hist[3 * k] = best;
    A.hist[3 * k + 1] = (double)bidx;
    A.hist[3 * k + 2] = (double)bsdt;
  }
  if (best < A.dt_min) {  // time_dr_advance.f90:122-127
    ts->err = 3; ts->done = 1; ts->skip = 1; ts->err_step = ts->nstep;
  }
}

// ---------------------------------------------------------------- ensembles
// Many independent instances stepped together: the mesh is the disjoint
// union of the instances' meshes, instance k owning elements [eoff[k],
// eoff[k+1]).  Every instance has at least kBlock elements, so a workgroup
// of F1 spans at most two instances (a and b); it keeps two partial slots of
// each control.  F1L gives each instance one workgroup per chunk (4096 unless tuned) of its
// elements, striding over the instance's own worklist segment, and E1b
// selects each instance's dt in one workgroup per instance.  Each instance's selection is the first-index
// extremum over its own (value, index) pairs, as a run of that instance
// alone makes it.

// E1's worklist appends into each instance's segment (one atomic per wave
// and instance) and its two partial slots of each control per workgroup
// (a workgroup spans at most two instances, a and b); every thread of the
// workgroup takes part
__device__ __forceinline__ void e1_collect(const FusedArgs& A, const EnsArgs& E, int e, int lane,
                                           bool moving, double cflv, double divv) {
  const int lb = xcd_block(), nb = gridDim.x;
  const int ia = A.einst[lb * kBlock];
  const int ib = A.einst[min(lb * kBlock + kBlock - 1, A.nel - 1)];
  const int my = e < A.nel ? (int)A.einst[e] : ib;
  if (!(cflv == cflv)) cflv = DBL_MAX;
  if (!(divv == divv)) divv = -1.0;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int k = s == 0 ? ia : ib;
    if (s == 1 && ib == ia) break;
    const uint64_t mask = __ballot(moving && my == k);
    if (mask) {
      const int lead = __ffsll((unsigned long long)mask) - 1;
      uint32_t base = 0;
      if (lane == lead) base = atomicAdd(E.cnt + k, (uint32_t)__popcll(mask));
      base = __shfl(base, lead, 64);
      const uint32_t slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
      if (moving && my == k && slot < (uint32_t)(E.eoff[k + 1] - E.eoff[k]))
        A.work[E.eoff[k] + slot] = e;
    }
  }
  block_reduce_at<false>(my == ia ? divv : -1.0, e + 1, E.part + lb, E.parti + lb);
  __syncthreads();
  block_reduce_at<false>(my == ib ? divv : -1.0, e + 1, E.part + nb + lb, E.parti + nb + lb);
  __syncthreads();
  block_reduce_at<true>(my == ia ? cflv : DBL_MAX, e + 1, E.part + 2 * nb + lb,
                        E.parti + 2 * nb + lb);
  __syncthreads();
  block_reduce_at<true>(my == ib ? cflv : DBL_MAX, e + 1, E.part + 3 * nb + lb,
                        E.parti + 3 * nb + lb);
}

// E1: F1 with per-instance partial slots and worklist segments
__global__ void __launch_bounds__(kBlock)
k_e1_getdt(FusedArgs A, FusedNodes N, EnsArgs E) {
  if (step_skipped(A, A.ts)) return;
  const int e = xcd_index();
  const int lane = threadIdx.x & 63;
  const uint64_t prev = A.rest[e >> 6];
  double cflv = DBL_MAX, divv = -1.0;
  bool moving = false, zeroed = false;
  if (e < A.nel) {
    int nd[4];
    load_nd(A.ielnd, e, nd);
    double x[4], y[4], u[4], v[4];
    load4(N.x, nd, x); load4(N.y, nd, y); load4(N.u, nd, u); load4(N.v, nd, v);
    Iso g;
    getiso(x, y, g);
    f1_terms(A, e, x, y, u, v, g, A.cs2[e], moving, cflv, divv);
    if (!moving && !A.fold && !((prev >> lane) & 1ull)) {
      const double z[4] = {0.0, 0.0, 0.0, 0.0};
      st4(A.cnviscx, e, z);
      st4(A.cnviscy, e, z);
      A.elvisc[e] = 0.0;
      zeroed = true;
    }
  }
  rest_update(A, e, lane, prev, e < A.nel && !moving, zeroed);
  e1_collect(A, E, e, lane, moving, cflv, divv);
}

// E1L: instance k's workgroups lboff[k] .. lboff[k+1]-1 stride over its
// worklist segment
__global__ void __launch_bounds__(kBlock)
k_e1_limiter(FusedArgs A, FusedNodes N, EnsArgs E) {
  if (step_skipped(A, A.ts)) return;
  int lo = 0, hi = E.ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (E.lboff[mid] <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const int k = lo, sub = blockIdx.x - E.lboff[k], nbk = E.lboff[k + 1] - E.lboff[k];
  const int base = E.eoff[k];
  const uint32_t count = min(E.cnt[k], (uint32_t)(E.eoff[k + 1] - base));
  double best = DBL_MAX;
  int bi = 0x7fffffff;
  for (uint32_t q = sub * blockDim.x + threadIdx.x; q < count; q += nbk * blockDim.x) {
    const int e = A.work[base + q];
    better_min(best, bi, f1_moving(A, N, e), e + 1);
  }
  block_reduce_at<true>(best, bi, E.lpart + blockIdx.x, E.lparti + blockIdx.x);
}

// E1b: one workgroup per instance -- its partials, then its dt selection;
// workgroup 0 also runs the step prologue of the whole ensemble.  No
// workgroup writes ts->done, which every workgroup reads at its start: an
// instance whose selection fails lowers ts->epend instead (the smallest
// failing instance wins, whatever the workgroup order), the kernels after
// this one skip the step on it, and the step's last kernel sets done.
__global__ void __launch_bounds__(kBlock) k_e1b_select(FusedArgs A, EnsArgs E, int nb) {
  DevTimestep* ts = A.ts;
  const int k = blockIdx.x, t = threadIdx.x;
  const bool skip = step_skipped(A, ts);
  __syncthreads();
  if (t == 0) {
    if (k == 0) {
      ts->skip = skip ? 1 : 0;
      A.dtres->negvol = 0;
      A.dtres->negvol2 = 0;
      if (!skip) {
        ts->nstep = ts->nstep + 1;
        ts->last = ts->nstep == ts->last_step ? 1 : 0;  // the tile step's last
      }
    }
    if (!skip)
      atomicAdd(A.vstats, (unsigned long long)min(E.cnt[k], (uint32_t)(E.eoff[k + 1] - E.eoff[k])));
    E.cnt[k] = 0;  // E1L has read it
  }
  if (skip) return;
  double va = -1.0, vc = DBL_MAX;
  int xa = 0x7fffffff, xc = 0x7fffffff;
  if (E.toff) {
    // the tile step's slots: one per tile (nb tiles), the instance's own
    for (int s = E.toff[k] + t; s < E.toff[k + 1]; s += kBlock) {
      better_max(va, xa, A.part[s], A.parti[s]);
      better_min(vc, xc, A.part[2 * nb + s], A.parti[2 * nb + s]);
    }
  } else {
    const int lb0 = E.eoff[k] / kBlock, lb1 = (E.eoff[k + 1] - 1) / kBlock;
    for (int lb = lb0 + t; lb <= lb1; lb += kBlock) {
      const int slot = (int)A.einst[lb * kBlock] == k ? 0 : 1;
      better_max(va, xa, E.part[slot * nb + lb], E.parti[slot * nb + lb]);
      better_min(vc, xc, E.part[(2 + slot) * nb + lb], E.parti[(2 + slot) * nb + lb]);
    }
  }
  for (int b = E.lboff[k] + t; b < E.lboff[k + 1]; b += kBlock)
    better_min(vc, xc, E.lpart[b], E.lparti[b]);
  // thread 0 writes both results and is the only reader
  double sva, svc;
  int32_t sxa, sxc;
  block_reduce_at<false>(va, xa, &sva, &sxa);
  __syncthreads();
  block_reduce_at<true>(vc, xc, &svc, &sxc);
  if (t != 0) return;
  double div_max;
  int div_idx;
  if (sva > DBL_MIN) { div_max = sva; div_idx = sxa; }
  else { div_max = DBL_MIN; div_idx = 0; }
  InstTs& it = E.inst[k];
  const int off = E.eoff[k];
  if (svc < 0.0) {  // hydro_kn_getdt.f90:75-78
    atomicMin(&ts->epend, 4 * k + 2);
    return;
  }
  double best;
  int bidx, bsdt;
  select_dt(A, it.dt, it.nstep, svc, sxc - off, div_max, div_idx > 0 ? div_idx - off : 0,
            best, bidx, bsdt);
  const int hstep = it.nstep - ts->hist_base;
  if (E.hist && hstep >= 0 && hstep < E.hist_cap) {
    double* hrec = E.hist + 3 * ((size_t)hstep * E.ninst + k);
    hrec[0] = best;
    hrec[1] = (double)bidx;
    hrec[2] = (double)bsdt;
  }
  it.nstep = it.nstep + 1;
  it.dt = best;
  it.idtel = bidx;
  it.sdt = bsdt;
  it.time = it.time + best;
  E.inst_dt[k] = best;
  E.inst_rdt[k] = 1.0 / best;
  if (best < A.dt_min) atomicMin(&ts->epend, 4 * k + 3);  // time_dr_advance.f90:122-127
}

// an ensemble instance's dt selection failed in this step (E1b): nothing
// after the selection runs, as the reference halts in time_dr_end
__device__ __forceinline__ bool ens_pending(const FusedArgs& A, const DevTimestep* ts) {
  return A.einst != nullptr && ts->epend != kNoPending;
}

// component terms of a mixed cell in F2 (hydro_dr_getforce.f90:52-66,
// hydro_dr_getenergy.f90:47-62, geometry_dr_volume.f90:66-79,
// density_dr_getdensity.f90:40-43, eos_dr_geteos.f90:50-55,
// hydro_dr_getacceleration.f90:69-72): predictor component force at x^n,
// predictor component energy (cpu = frvolume*u^n corners, dts = dt/2),
// half-step component volume, density and pressure, corrector component
// force (kept in rcpscratch21/22 for F4), and the cell's corrector corner
// force replaced by the frvolume average of the component forces.
// Component energies are not saved and restored between predictor and
// corrector (hydro_dr_set.f90 copies elenergy only).
__device__ __forceinline__ void mix_f2(const FusedArgs& A, int imx, const Iso& g,
                                      const Iso& gh, const double u[4], const double v[4],
                                      double h, double fx[4], double fy[4]) {
  const MixArgs& M = A.mx;
  int icp = M.imxfcp[imx] - 1;
  const int n = M.imxncp[imx];
  double w1[4] = {0.0, 0.0, 0.0, 0.0}, w2[4] = {0.0, 0.0, 0.0, 0.0};
  for (int ii = 0; ii < n; ++ii, ++icp) {
    const double f = M.frvolume[icp], cm = M.cpmass[icp];
    double cfx[4], cfy[4], cu[4], cv[4];
    mix_force(M, icp, M.cppressure[icp], g.a1, g.a3, g.b1, g.b3, cfx, cfy);
#pragma unroll
    for (int j = 0; j < 4; ++j) { cu[j] = f * u[j]; cv[j] = f * v[j]; }
    const double ce = energy_update(M.cpenergy[icp], cm, cfx, cfy, cu, cv, h, A.zerocut);
    M.cpenergy[icp] = ce;
    const double crho = cm / (gh.vol * f);
    int cmat = M.icpmat[icp];
    cmat = cmat > 0 ? cmat : 0;
    const double cph = eos_pressure(A.eos_itype[cmat], A.eos_param + BL_NEOSPARAM * cmat, crho,
                                    ce, A.pcut);
    mix_force(M, icp, cph, gh.a1, gh.a3, gh.b1, gh.b3, cfx, cfy);
    st4c(M.cpfx, icp, cfx);
    st4c(M.cpfy, icp, cfy);
#pragma unroll
    for (int j = 0; j < 4; ++j) { w1[j] = w1[j] + f * cfx[j]; w2[j] = w2[j] + f * cfy[j]; }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) { fx[j] = w1[j]; fy[j] = w2[j]; }
}

// component terms of a mixed cell in F4: corrector geometry (the cell's iso
// terms at x^{n+1}), volume, density, energy with cpu = frvolume*ubar
// corners, EOS
__device__ __forceinline__ void mix_f4(const FusedArgs& A, int imx, const Iso& g,
                                      const double ub[4], const double vb[4], double dt) {
  const MixArgs& M = A.mx;
  int icp = M.imxfcp[imx] - 1;
  const int n = M.imxncp[imx];
  for (int ii = 0; ii < n; ++ii, ++icp) {
    const double f = M.frvolume[icp], cm = M.cpmass[icp];
    M.cpa1[icp] = g.a1; M.cpa3[icp] = g.a3; M.cpb1[icp] = g.b1; M.cpb3[icp] = g.b3;
    const double cvol = g.vol * f;
    M.cpvolume[icp] = cvol;
    const double crho = cm / cvol;
    M.cpdensity[icp] = crho;
    double cfx[4], cfy[4], cu[4], cv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { cu[j] = f * ub[j]; cv[j] = f * vb[j]; }
    ld4c(M.cpfx, icp, cfx);
    ld4c(M.cpfy, icp, cfy);
    const double ce = energy_update(M.cpenergy[icp], cm, cfx, cfy, cu, cv, dt, A.zerocut);
    M.cpenergy[icp] = ce;
    int cmat = M.icpmat[icp];
    cmat = cmat > 0 ? cmat : 0;
    const int et = A.eos_itype[cmat];
    const double* ep = A.eos_param + BL_NEOSPARAM * cmat;
    M.cppressure[icp] = eos_pressure(et, ep, crho, ce, A.pcut);
    M.cpcs2[icp] = eos_cs2(et, ep, crho, ce, A.pcut, A.ccut);
  }
}

// F2m, in the workgroups F2 appends after its elements: one thread per
// mixed cell; F2 leaves the cell's corrector corner force to it.  It
// recomputes the cell's geometry at x^n and at the half step exactly as F2
// does, then the component terms, whose frvolume average is the cell's
// force.  Nothing else F2 reads or writes is touched, so the two run side by
// side.
__device__ __forceinline__ void f2_mix_cell(const FusedArgs& A, const FusedNodes& N,
                                            const DevTimestep* ts, int imx) {
  const int e = A.mx.imxel[imx] - 1;
  const double h = 0.5 * ts->dt;
  int nd[4];
  load_nd(A.ielnd, e, nd);
  double x[4], y[4], u[4], v[4], xh[4], yh[4];
  load4(N.x, nd, x); load4(N.y, nd, y); load4(N.u, nd, u); load4(N.v, nd, v);
#pragma unroll
  for (int c = 0; c < 4; ++c) { xh[c] = x[c] + h * u[c]; yh[c] = y[c] + h * v[c]; }
  Iso g, gh;
  getiso(x, y, g);
  getiso(xh, yh, gh);
  double fx[4], fy[4];
  mix_f2(A, imx, g, gh, u, v, h, fx, fy);
  const int nel = A.nel;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    A.fxT[j * nel + e] = fx[j];
    A.fyT[j * nel + e] = fy[j];
  }
}

// F4m, in the workgroups F5 appends after its first: the components'
// corrector state of a mixed cell (F4 writes only the cell's own fields,
// which this does not read)
__device__ __forceinline__ void f4_mix_cell(const FusedArgs& A, const FusedNodes& O,
                                            const DevTimestep* ts, int imx) {
  const int e = A.mx.imxel[imx] - 1;
  int nd[4];
  load_nd(A.ielnd, e, nd);
  double x[4], y[4], ub[4], vb[4];
  load4(O.x, nd, x); load4(O.y, nd, y); load4(A.ubar, nd, ub); load4(A.vbar, nd, vb);
  Iso g;
  getiso(x, y, g);
  mix_f4(A, imx, g, ub, vb, ts->dt);
}

// ---------------------------------------------------------------- F2
// F2's work for element e: the predictor at x^n (force with p^n, half-step
// geometry, density, energy and EOS in registers) and the corrector's corner
// forces at the half step P+Q(+SP)(+HG) in fx, fy, with the half-step
// corner weights wt and each corner's nodal-mass term mt as
// hydro_kn_scatteracceleration picks it (hydro_kn_getacceleration.f90:58-67).
// at_rest: the element's viscosity forces are +0.0 and are not read.  With
// check_volume the element reports a negative half-step volume.  (A mixed
// cell's corner force comes from its components, F2m.)
__device__ __forceinline__ void f2_element(const FusedArgs& A, const FusedNodes& N, int e,
                                           const int nd[4], double dt, double rdt,
                                           bool at_rest, bool check_volume, double fx[4],
                                           double fy[4], double wt[4], double mt[4]) {
  const double h = 0.5 * dt;  // geometry_dr_getvertex: 0.5*dt
  // the element's inputs first, in one group of loads before any branch
  // (the compiler waits for every load in flight where branches join); the
  // region index is a load either way, used with reg_tables_force only
  double x[4], y[4], u[4], v[4];
  load4(N.x, nd, x); load4(N.y, nd, y); load4(N.u, nd, u); load4(N.v, nd, v);
  const double pre = A.pre[e], m = A.mass[e], en0 = A.en[e];
  int mat = A.ielmat[e];
  const int ireg_raw = (A.reg_tables_force ? A.ielreg : A.ielmat)[e];
  double qx[4], qy[4];
  if (at_rest) {
#pragma unroll
    for (int c = 0; c < 4; ++c) { qx[c] = 0.0; qy[c] = 0.0; }
  } else {
    ld4(A.cnviscx, e, qx);
    ld4(A.cnviscy, e, qy);
  }
  mat = mat > 0 ? mat : 0;
  const int et = A.eos_itype[mat];
  const double* ep = A.eos_param + BL_NEOSPARAM * mat;
  const int ireg = A.reg_tables_force ? ireg_raw - 1 : 0;
  // the hourglass coefficient and the EOS's first parameter, loaded ahead
  // of their use (each a load either way)
  const double kap = (A.zhg ? A.kappareg : A.eos_param)[A.zhg ? ireg : 0];
  const double ep0 = ep[0];
  // --- predictor: hydro_dr_getforce at x^n with p^n
  Iso g;
  getiso(x, y, g);
  forceP(pre, g.a1, g.a3, g.b1, g.b3, fx, fy);
#pragma unroll
  for (int c = 0; c < 4; ++c) { fx[c] = fx[c] + qx[c]; fy[c] = fy[c] + qy[c]; }
  if (A.zsp) {
    double spm[4];
    ld4(A.spmass, e, spm);
    // elcs2 at this point holds cs2 + elvisc (hydro_dr_getcs2 in getdt)
    forceSP(A.pmeritreg[ireg], A.rho[e], A.cs2[e] + (at_rest ? 0.0 : A.elvisc[e]), x, y, spm,
            fx, fy);
  }
  // --- predictor getstate: vertex, geometry, density; energy; EOS
  double xh[4], yh[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) { xh[c] = x[c] + h * u[c]; yh[c] = y[c] + h * v[c]; }
  Iso gh;
  getiso(xh, yh, gh);
  const double rhoh = m / gh.vol;
  const double eh = energy_update(en0, m, fx, fy, u, v, h, A.zerocut);
  const double ph = eos_pressure(et, ep0, ep, rhoh, eh, A.pcut);
  // the nodal masses, loaded here so that the corrector hides their latency
  double cm[4];
  ld4(A.cnmass, e, cm);
  // --- corrector force at the half step (zcorrector: HG on)
  forceP(ph, gh.a1, gh.a3, gh.b1, gh.b3, fx, fy);
#pragma unroll
  for (int c = 0; c < 4; ++c) { fx[c] = fx[c] + qx[c]; fy[c] = fy[c] + qy[c]; }
  if (A.zsp) {
    double spm[4];
    ld4(A.spmass, e, spm);
    const double cs2h = eos_cs2(et, ep, rhoh, eh, A.pcut, A.ccut);
    forceSP(A.pmeritreg[ireg], rhoh, cs2h, xh, yh, spm, fx, fy);
  }
  if (A.zhg) forceHG(kap, rhoh, gh.vol, rdt, u, v, fx, fy);
  // nodal-mass term of each corner: cnmass, else the previous corner's
  // cnmass, else rho*cnwt at the half step
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (cm[j] > A.zerocut) mt[j] = cm[j];
    else if (cm[(j + 3) & 3] > A.zerocut) mt[j] = cm[(j + 3) & 3];
    else mt[j] = rhoh * gh.wt[j];
    wt[j] = gh.wt[j];
  }
  // last: no load of the element passes this atomic
  if (check_volume && gh.vol < 0.0) atomicMax(&A.dtres->negvol, e + 1);
}

// the zero viscosity of an element at rest now but not in the step before
// (or on a run call's first step), written after the element's loads
__device__ __forceinline__ void write_rest_zeros(const FusedArgs& A, int e) {
  const double z[4] = {0.0, 0.0, 0.0, 0.0};
  st4(A.cnviscx, e, z);
  st4(A.cnviscy, e, z);
  A.elvisc[e] = 0.0;
}

// (three waves per SIMD: at most 168 VGPRs)
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_f2_force(FusedArgs A, FusedNodes N, int nbm) {
  const DevTimestep* ts = A.ts;
  if (ts->skip || ens_pending(A, ts)) return;
  // the mixed cells' workgroups come first, so that their longer serial
  // work starts with the grid rather than trailing it; nbm is a multiple of
  // the XCD count, so the element workgroups keep their XCD mapping
  if ((int)blockIdx.x < nbm) {
    const int imx = (int)blockIdx.x * kBlock + threadIdx.x;
    if (imx < A.mx.nmx) f2_mix_cell(A, N, ts, imx);
    return;
  }
  const int e = xcd_block_of((int)blockIdx.x - nbm, (int)gridDim.x - nbm) * kBlock + threadIdx.x;
  if (e >= A.nel) return;
  const int inst = A.einst ? (int)A.einst[e] : -1;
  const double dt = inst >= 0 ? A.inst_dt[inst] : ts->dt;
  const double rdt = inst >= 0 ? A.inst_rdt[inst] : ts->rdt;
  // an element F1 found at rest has +0.0 viscosity forces: not read.  With
  // F1 folded into F4, F2 writes those zeros when the element moved in the
  // step before (or this is the first step of the run call)
  const bool at_rest = (A.rest[e >> 6] >> (e & 63)) & 1ull;
  int nd[4];
  load_nd(A.ielnd, e, nd);
  double fx[4], fy[4], wt[4], mt[4];
  f2_element(A, N, e, nd, dt, rdt, at_rest, true, fx, fy, wt, mt);
  const int nel = A.nel;
  // a mixed cell's force comes from its components (F2m)
  const bool own_force = !(A.mx.ncp > 0 && A.ielmat[e] < 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (own_force) {
      A.fxT[j * nel + e] = fx[j];
      A.fyT[j * nel + e] = fy[j];
    }
    A.wtT[j * nel + e] = wt[j];
    A.mT[j * nel + e] = mt[j];
  }
  if (A.fold) {
    // the zeros (last, so that no load of this kernel waits behind them)
    const bool zeroed = at_rest && !((A.rest_prev[e >> 6] >> (e & 63)) & 1ull);
    if (zeroed) write_rest_zeros(A, e);
    // lanes past nel left at the start, and they are the wave's last lanes
    const uint64_t wrote = __ballot(zeroed);
    if ((threadIdx.x & 63) == 0 && wrote)
      count_zeroed(A, wrote);
  }
}

// ---------------------------------------------------------------- F3
// F3's work for a node from its sums over its corners (mass m, area a,
// force fu, fv in the (corner, element) order of the reference's scatter),
// its position and velocity at step start and its type t:
// hydro_kn_getacceleration, utils_kn_bc, hydro_kn_applyacceleration and the
// corrector's geometry_kn_getvertex.  out = x^{n+1}, y^{n+1}, u^{n+1},
// v^{n+1}, ubar, vbar.
__device__ __forceinline__ void f3_node(const FusedArgs& A, double dt, double m, double a,
                                        double fu, double fv, double x0, double y0, double u0,
                                        double v0, int t, double out[6]) {
  const double h = 0.5 * dt;
  const double w1 = A.dencut * a;
  if (m > w1) { fu = fu / m; fv = fv / m; }
  else { fu = 0.0; fv = 0.0; }
  if (t == -1) fu = 0.0;
  else if (t == -2) fv = 0.0;
  else if (t == -3) { fu = 0.0; fv = 0.0; }
  if (fu * fu + fv * fv < A.accut2) { fu = 0.0; fv = 0.0; }
  const double u1 = u0 + dt * fu, v1 = v0 + dt * fv;
  const double dt05 = 0.5 * dt;
  out[4] = u0 + dt05 * fu;
  out[5] = v0 + dt05 * fv;
  out[2] = u1;
  out[3] = v1;
  const double xh = x0 + h * u0, yh = y0 + h * v0;
  out[0] = xh + h * u1;
  out[1] = yh + h * v1;
}

__global__ void __launch_bounds__(kBlock)
k_f3_accel(FusedArgs A, FusedNodes N, FusedNodesOut O) {
  const DevTimestep* ts = A.ts;
  if (ts->skip || A.dtres->negvol || ens_pending(A, ts)) return;
  const int n = xcd_index();
  if (n >= A.nnd) return;
  const double dt = A.ninst ? A.inst_dt[A.ninst[n]] : ts->dt;
  double m = 0.0, a = 0.0, fu = 0.0, fv = 0.0;
  const int b = A.hyd_ptr[n], end = A.hyd_ptr[n + 1];
  const int nel = A.nel;
  // incidences in (corner j, element) order; corner-major terms make the
  // loads of a wave for one corner slot contiguous
  for (int k = b; k < end; ++k) {
    const int c = A.hyd_idx[k];
    const int off = (c & 3) * nel + (c >> 2);
    m = m + A.mT[off];
    a = a + A.wtT[off];
    fu = fu + A.fxT[off];
    fv = fv + A.fyT[off];
  }
  double out[6];
  f3_node(A, dt, m, a, fu, fv, N.x[n], N.y[n], N.u[n], N.v[n], A.indtype[n], out);
  O.x[n] = out[0];
  O.y[n] = out[1];
  O.u[n] = out[2];
  O.v[n] = out[3];
  A.ubar[n] = out[4];
  A.vbar[n] = out[5];
}

// F4's work for element e from its nodes at x^{n+1} (x, y), their ubar
// (ub, vb) and its corrector corner forces (fx, fy): geometry g, density,
// energy with ubar, EOS, written to the output element fields (en_o ...,
// the same arrays as en ... unless the step writes the other element set);
// returns the new cs2
// the element inputs of F4: mass, energy at the step's start, EOS type and
// parameters of its material (eos_param(:, max(mat, 0)))
struct ElIn {
  double m, en;
  int et;
  double p[BL_NEOSPARAM];
};
// the EOS of material mat (ielmat; 0 for a value <= 0)
__device__ __forceinline__ void el_eos(const FusedArgs& A, int mat, ElIn& r) {
  mat = mat > 0 ? mat : 0;
  r.et = A.eos_itype[mat];
#pragma unroll
  for (int k = 0; k < BL_NEOSPARAM; ++k) r.p[k] = A.eos_param[BL_NEOSPARAM * mat + k];
}

// F4 with its element inputs m, en0 (energy at the step's start), the EOS
// type et and parameters ep
__device__ __forceinline__ double f4_state_from(const FusedArgs& A, int e, double dt, double m,
                                                double en0, int et, const double* ep,
                                                const double x[4], const double y[4],
                                                const double ub[4], const double vb[4],
                                                const double fx[4], const double fy[4], Iso& g) {
  getiso(x, y, g);
  const double rho = m / g.vol;
  const double en = energy_update(en0, m, fx, fy, ub, vb, dt, A.zerocut);
  A.en_o[e] = en;
  A.rho_o[e] = rho;
  A.pre_o[e] = eos_pressure(et, ep, rho, en, A.pcut);
  const double c2 = eos_cs2(et, ep, rho, en, A.pcut, A.ccut);
  A.cs2_o[e] = c2;
  if (A.write_geometry) {
    // the remap that follows reads the corner weights and the volume at
    // x^{n+1}; it rewrites the rest of the geometry on the moved mesh
    st4(A.cnwt, e, g.wt);
    A.vol[e] = g.vol;
  }
  // last: no load of the element passes this atomic
  if (g.vol < 0.0) atomicMax(&A.dtres->negvol2, e + 1);
  return c2;
}
__device__ __forceinline__ double f4_state_of(const FusedArgs& A, int e, double dt,
                                              const double x[4], const double y[4],
                                              const double ub[4], const double vb[4],
                                              const double fx[4], const double fy[4], Iso& g) {
  int mat = A.ielmat[e];
  mat = mat > 0 ? mat : 0;
  return f4_state_from(A, e, dt, A.mass[e], A.en[e], A.eos_itype[mat],
                       A.eos_param + BL_NEOSPARAM * mat, x, y, ub, vb, fx, fy, g);
}

// F4's work for element e with its inputs from F3's node set and F2's
// corner-major forces; leaves nd, x, y and the geometry g for F1's terms
__device__ __forceinline__ double f4_element(const FusedArgs& A, const FusedNodes& O,
                                             const DevTimestep* ts, int e, int nd[4],
                                             double x[4], double y[4], Iso& g) {
  const double dt = A.einst ? A.inst_dt[A.einst[e]] : ts->dt;
  load_nd(A.ielnd, e, nd);
  double ub[4], vb[4];
  load4(O.x, nd, x); load4(O.y, nd, y); load4(A.ubar, nd, ub); load4(A.vbar, nd, vb);
  double fx[4], fy[4];
  const int nel = A.nel;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    // read once in the step (k_materialize reads the last step's copy
    // again after the batch): non-temporal loads, F4 2 µs faster at 1 M
    // elements
    fx[j] = __builtin_nontemporal_load(A.fxT + j * nel + e);
    fy[j] = __builtin_nontemporal_load(A.fyT + j * nel + e);
  }
  return f4_state_of(A, e, dt, x, y, ub, vb, fx, fy, g);
}

// ---------------------------------------------------------------- F4
__global__ void __launch_bounds__(kBlock)
k_f4_state(FusedArgs A, FusedNodes O) {
  const DevTimestep* ts = A.ts;
  if (ts->skip || A.dtres->negvol) return;
  const int e = xcd_index();
  if (!A.fold) {
    if (e >= A.nel) return;
    int nd[4];
    double x[4], y[4];
    Iso g;
    f4_element(A, O, ts, e, nd, x, y, g);
    return;
  }
  // F1 of the next step, from x^{n+1}, u^{n+1} and the cs2 just computed
  const int lane = threadIdx.x & 63;
  double cflv = DBL_MAX, divv = -1.0;
  bool moving = false;
  if (e < A.nel) {
    int nd[4];
    double x[4], y[4], u[4], v[4];
    Iso g;
    const double c2 = f4_element(A, O, ts, e, nd, x, y, g);
    load4(O.u, nd, u); load4(O.v, nd, v);
    f1_terms(A, e, x, y, u, v, g, c2, moving, cflv, divv);
  }
  const uint64_t now = __ballot(e < A.nel && !moving);
  if (lane == 0) A.rest_next[e >> 6] = now;
  f1_collect(A, e, lane, moving, cflv, divv);
}

// F4 of an ensemble step with E1 of the next step folded in (E1's terms as
// k_f4_state's fold, E1's per-instance worklist segments and partial slots)
__global__ void __launch_bounds__(kBlock)
k_f4_state_ens(FusedArgs A, FusedNodes O, EnsArgs E) {
  const DevTimestep* ts = A.ts;
  if (ts->skip || A.dtres->negvol || ens_pending(A, ts)) return;
  const int e = xcd_index();
  const int lane = threadIdx.x & 63;
  double cflv = DBL_MAX, divv = -1.0;
  bool moving = false;
  if (e < A.nel) {
    int nd[4];
    double x[4], y[4], u[4], v[4];
    Iso g;
    const double c2 = f4_element(A, O, ts, e, nd, x, y, g);
    load4(O.u, nd, u); load4(O.v, nd, v);
    f1_terms(A, e, x, y, u, v, g, c2, moving, cflv, divv);
  }
  const uint64_t now = __ballot(e < A.nel && !moving);
  if (lane == 0) A.rest_next[e >> 6] = now;
  e1_collect(A, E, e, lane, moving, cflv, divv);
}


// ------------------------------------------------------------ the tile step
// F2, F3 and F4 of a pure-cell Lagrangian step in one kernel per tile of
// elements (bl_tiles.hip): F2 of the tile's own elements and of the ring of
// elements around their nodes, whose corner terms stay in LDS for the sums
// of every node of the own elements; the node results stay in LDS for the
// corrector state of the own elements, which the tile completes.  A node
// shared by several tiles is computed by each (the same inputs, the same
// bits) and written by one.  The arithmetic is F2's, F3's and F4's own
// (f2_element, f3_node, f4_state_of), so the results are the same bits.  A
// node's corner terms are summed over jj = 1..4 in that order: on a mesh
// where no node is corner jj of two elements (build_tiles checks it) this is
// the (corner, element) order of hydro_kn_scatteracceleration.  The element
// state of the step is written to the other element set, so a step that
// stops on a negative volume leaves the state it started from, as the
// three-kernel step does (F4 does not run after F2's negative volume there).

// An ensemble tile's elements in motion to its instance k's worklist
// segment, one atomic per wave (as e1_collect lists them); every lane of
// the wave takes part
__device__ __forceinline__ void ens_list(const FusedArgs& A, const EnsArgs& E, int k, int lane,
                                         bool moving, int e) {
  const uint64_t mask = __ballot(moving);
  if (!mask) return;
  const int lead = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == lead) base = atomicAdd(E.cnt + k, (uint32_t)__popcll(mask));
  base = __shfl(base, lead, 64);
  const uint32_t slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  if (moving && slot < (uint32_t)(E.eoff[k + 1] - E.eoff[k])) A.work[E.eoff[k] + slot] = e;
}

// The partial slots of one workgroup of F1's terms: its DIV maximum and
// CFL minimum to slots seg and 2 * nseg + seg.  Inactive lanes carry the
// neutral values and an index that never wins a tie.
__device__ __forceinline__ void seg_collect(const FusedArgs& A, int seg, int nseg, bool active,
                                            int e, double cflv, double divv) {
  const int idx = active ? e + 1 : 0x7fffffff;
  if (!active || !(divv == divv)) divv = -1.0;      // a NaN never wins the strict '>' scan
  if (!active || !(cflv == cflv)) cflv = DBL_MAX;   // nor the strict '<' scan
  block_reduce_div_cfl(divv, idx, cflv, idx, A.part + seg, A.parti + seg, A.part + 2 * nseg + seg,
                       A.parti + 2 * nseg + seg);
}

// (three waves per SIMD: at most 168 VGPRs; LDS allows three workgroups)
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_lag_tile(FusedArgs A, FusedNodes N, FusedNodesOut O, EnsArgs E) {
  // slot-fastest, so that a wave's lanes of one (quantity, corner) are
  // consecutive doubles: (quantity m, area, fx, fy; corner jj; node slot)
  __shared__ double acc[4][4][kTileNodes + 1];
  __shared__ double nout[6][kTileNodes + 1];  // x, y, u, v at n+1, ubar, vbar
  const int t = xcd_block();
  const int tid = threadIdx.x;
  // the tile's bounds: one scalar load
  const int4 tb = A.tile_bounds[t];
  const int eb = tb.x, ne = tb.y - tb.x;
  const int nb = tb.z, nn = tb.w - tb.z;
  // every thread loads a node entry and an element entry (a tile has at
  // least one of each; threads past the end reload the last), so that these
  // loads are in flight together; no barrier comes before phase A
  const bool own = tid < nn, active = tid < ne;
  const uint32_t nv = (uint32_t)A.tile_nodes[nb + min(tid, nn - 1)];
  // the node's corner mask: the slots of F2 that phase B sums
  const uint32_t cm = A.tile_ncm[nb + min(tid, nn - 1)];
  const int2 ent = A.tile_el[eb + min(tid, ne - 1)];
  const int4 q = A.tile_nd[eb + min(tid, ne - 1)];
  const DevTimestep* ts = A.ts;
  if (ts->skip || ens_pending(A, ts)) return;
  // the high bit of the node's entry: this tile writes it (other tiles
  // holding it compute the same values for their own elements)
  const int n = (int)(nv & 0x7fffffffu);
  const bool writes = own && (nv >> 31) != 0;
  // an ensemble tile lies in one instance: its dt
  int inst = 0;
  double dt = ts->dt, rdt = ts->rdt;
  if (A.einst) {
    inst = A.einst[A.tile_el[eb].x & 0x3fffffff];
    dt = A.inst_dt[inst];
    rdt = A.inst_rdt[inst];
  }
  // --- F2: every element around the tile's nodes, one per thread
  // (every thread's entry names an element of the tile)
  const int e = ent.x & 0x3fffffff;
  const uint32_t flags = active ? (uint32_t)ent.x >> 30 : 0u;
  const uint32_t slots = (uint32_t)ent.y;
  // its at-rest bytes now and in the step before, and whether this is the
  // batch's last step, read before any branch
  const bool at_rest = A.rest8[e] != 0, rest_before = A.rest8_prev[e] != 0;
  const bool last = ts->last != 0;
  double fx[4], fy[4];
  bool zeroed = false;
  if (active) {
    const int nd[4] = {q.x, q.y, q.z, q.w};
    const bool writer = (flags & kTileOwn) != 0;
    double wt[4], mt[4];
    f2_element(A, N, e, nd, dt, rdt, at_rest, writer, fx, fy, wt, mt);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t sl = (slots >> (8 * j)) & 0xFFu;
      if (sl != kNoSlot) {
        acc[0][j][sl] = mt[j];
        acc[1][j][sl] = wt[j];
        acc[2][j][sl] = fx[j];
        acc[3][j][sl] = fy[j];
      }
    }
    if (writer) {
      // the batch's last step leaves the corner forces in rscratch23/24
      // (lagstep cnfx/cnfy)
      if (last) {
        st4_nt(A.cnfx, e, fx);
        st4_nt(A.cnfy, e, fy);
      }
      zeroed = at_rest && !rest_before;
      if (zeroed) write_rest_zeros(A, e);
    }
  }
  const uint64_t wrote = __ballot(zeroed);
  if ((tid & 63) == 0 && wrote) count_zeroed(A, wrote);
  // the tile's node's start-of-step values for F3 (L2 hits: phase A's
  // elements read them), loaded after phase A's work to spare its registers
  // and before the barrier, whose wait hides their latency
  const double x0 = N.x[n], y0 = N.y[n], u0 = N.u[n], v0 = N.v[n];
  const int ntype = A.indtype[n];
  __syncthreads();
  // phase C's element inputs, loaded before phase B so that their latency
  // overlaps it (the tables they index after it); every thread loads them
  // (a ring element's are L2 hits), as loads under a branch make the
  // compiler wait for them where the branches join
  const bool interior = active && (flags & kTileOwn);
  ElIn ein;
  ein.m = A.mass[e];
  ein.en = A.en[e];
  const int emat = A.ielmat[e];
  // (used with reg_tables only)
  const int ereg = (A.reg_tables ? A.ielreg : A.ielmat)[e];
  // the EOS and region tables of those (their loads overlap F3; the index
  // loads arrive with the node values F3 waits for)
  el_eos(A, emat, ein);
  const CflIn cin = cfl_of_region(A, ereg - 1);
  // --- F3: the tile's nodes, their corners in jj order
  if (own) {
    double m = 0.0, a = 0.0, fu = 0.0, fv = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((cm >> j) & 1u) {
        m = m + acc[0][j][tid];
        a = a + acc[1][j][tid];
        fu = fu + acc[2][j][tid];
        fv = fv + acc[3][j][tid];
      }
    double out[6];
    f3_node(A, dt, m, a, fu, fv, x0, y0, u0, v0, ntype, out);
    if (writes) {
      O.x[n] = out[0];
      O.y[n] = out[1];
      O.u[n] = out[2];
      O.v[n] = out[3];
      A.ubar[n] = out[4];
      A.vbar[n] = out[5];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) nout[k][tid] = out[k];
  }
  __syncthreads();
  // --- F4 and the next step's F1 terms: the elements held whole
  double cflv = DBL_MAX, divv = -1.0;
  bool moving = false;
  if (interior) {
    double x[4], y[4], u[4], v[4], ub[4], vb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t sl = (slots >> (8 * j)) & 0xFFu;
      x[j] = nout[0][sl];
      y[j] = nout[1][sl];
      u[j] = nout[2][sl];
      v[j] = nout[3][sl];
      ub[j] = nout[4][sl];
      vb[j] = nout[5][sl];
    }
    Iso g;
    const double c2 =
        f4_state_from(A, e, dt, ein.m, ein.en, ein.et, ein.p, x, y, ub, vb, fx, fy, g);
    f1_terms(A, e, x, y, u, v, g, c2, cin, moving, cflv, divv);
    A.rest8_next[e] = moving ? 0 : 1;
  }
  // an ensemble's next-step limiter reads its instance's worklist
  if (A.einst) ens_list(A, E, inst, tid & 63, interior && moving, e);
  seg_collect(A, t, A.nseg, interior, e, cflv, divv);
}

// F1 of a tile-step run call's first step, in the tile step's segments:
// one workgroup per tile over its own elements; the at-rest bytes of that
// step go to rest8_next
__global__ void __launch_bounds__(kBlock)
k_f1_seg(FusedArgs A, FusedNodes N, EnsArgs E) {
  if (step_skipped(A, A.ts)) return;
  const int tid = threadIdx.x;
  const int b = xcd_block();
  const int eb = A.tile_eptr[b], ne = A.tile_eptr[b + 1] - eb;
  const int2 ent = tid < ne ? A.tile_el[eb + tid] : make_int2(0, 0);
  const bool active = tid < ne && (((uint32_t)ent.x >> 30) & kTileOwn);
  const int e = ent.x & 0x3fffffff, seg = b;
  double cflv = DBL_MAX, divv = -1.0;
  bool moving = false;
  if (active) {
    int nd[4];
    load_nd(A.ielnd, e, nd);
    double x[4], y[4], u[4], v[4];
    load4(N.x, nd, x); load4(N.y, nd, y); load4(N.u, nd, u); load4(N.v, nd, v);
    Iso g;
    getiso(x, y, g);
    f1_terms(A, e, x, y, u, v, g, A.cs2[e], moving, cflv, divv);
    A.rest8_next[e] = moving ? 0 : 1;
  }
  // an ensemble tile: its instance's worklist for the limiter pass
  if (A.einst)
    ens_list(A, E, A.einst[A.tile_el[eb].x & 0x3fffffff], tid & 63, active && moving, e);
  seg_collect(A, seg, A.nseg, active, e, cflv, divv);
}

// ------------------------------------------------------- remap (Eulerian)
// The unsplit remap of ale_dr_advect (adv_type = 1) with zeul or the node
// relaxation's mesh velocity, in five grid kernels.  Intermediates live in
// the handle's own buffers (ale_tmp) rather than the alestep scratch ids,
// which nothing reads after the step.  On a
// conforming mesh the neighbour's flux volume of a shared face is a function
// of the two shared nodes only, so each element computes it itself with the
// neighbour's argument order (bit-identical to the neighbour's own value).
struct AleArgs {
  int nel, nnd, zeul;
  double zerocut, dencut, pcut, ccut;
  const int32_t* ielnd;
  const int32_t* ielel;
  const int32_t* ielfc;
  const int32_t* nbfc;  // ielel and ielfc packed per face (load_nbfc_packed)
  const int32_t* indtype;
  const int32_t* ielmat;
  const int32_t* eos_itype;
  const double* eos_param;
  const int32_t* ale_ptr;  // node -> corners, ALE order (ielsort2, corner)
  const int32_t* ale_idx;
  const int32_t* el_ptr;   // node -> corners, element order
  const int32_t* el_idx;
  double* x;  // node set: x^{n+1}, u^{n+1} in, remapped out
  double* y;
  double* u;
  double* v;
  const double* ubar;
  const double* vbar;
  double *cnwt, *rho, *vol, *mass, *en, *pre, *cs2, *cnmass;
  double *cnx, *cny, *a1, *a2, *a3, *b1, *b2, *b3;
  // temporaries: (4, nel) face / corner arrays and (nel) scalars; the
  // arrays the node pass gathers (cnm0T, cnflux, fu, fv) are corner-major,
  // X[j*nel + e], so one corner slot of a wave's nodes is contiguous
  double *fcdm, *flux, *cnm0, *cndm, *cnflux, *fu, *fv;
  double *dv, *v0, *m0, *cnm0T;
  DtResult* dtres;
  const DevTimestep* ts;
};

#define REMAP_GATE(B) \
  if ((B).ts->skip || !(B).ts->ale_on) return

// nodes of e at x^{n+1} and after the mesh motion (ale_kn_getmeshvelocity
// + geometry_kn_getvertex, ale_dr_getfluxvolume.f90:49-59), and the flux
// volumes of e's faces (own) and of the neighbours' matching faces (nbv)
__device__ __forceinline__ void remap_faces(const AleArgs& B, const int nd[4], double dt,
                                            double ax[4], double ay[4], double bx[4],
                                            double by[4], double own[4], double nbv[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double wu = B.ubar[nd[j]], wv = B.vbar[nd[j]];
    if (B.zeul) { wu = -wu; wv = -wv; }
    ax[j] = B.x[nd[j]];
    ay[j] = B.y[nd[j]];
    bx[j] = ax[j] + dt * wu;
    by[j] = ay[j] + dt * wv;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int jp = (j + 1) & 3;
    own[j] = fluxvolume(ax[j], ay[j], ax[jp], ay[jp], bx[jp], by[jp], bx[j], by[j], B.zerocut);
    nbv[j] = fluxvolume(ax[jp], ay[jp], ax[j], ay[j], bx[j], by[j], bx[jp], by[jp], B.zerocut);
  }
}

// R1: ale_kn_sumflux(fcdv) -> volume change, ale_kn_fluxelvl(cnwt, fcdv,
// rho) -> face mass fluxes (ale_dr_advect.f90:175-190)
// With FOLD (Eulerian ALE: no kernel between F4 and R1 reads what F5
// writes; the mixed cells' F4m then runs in R4), R1 also does F5's work: workgroup 0's first thread
// records the step's negative-volume outcome in ts, and every workgroup
// tests the negative-volume flags itself rather than the ts->skip that
// thread may be setting.
template <bool FOLD>
__global__ void __launch_bounds__(kBlock) k_r1(AleArgs B, DevTimestep* tsw) {
  if (FOLD) {
    const DtResult* res = B.dtres;
    const bool skip0 = tsw->skip != 0;  // as F1b left it
    if (blockIdx.x == 0 && threadIdx.x == 0 && !skip0) {
      if (res->negvol) {
        tsw->err = 1; tsw->err_idx = -res->negvol; tsw->done = 1; tsw->skip = 1;
        tsw->err_step = tsw->nstep;
      } else {
        tsw->flips = tsw->flips + 1;
        if (res->negvol2) {
          tsw->err = 1; tsw->err_idx = -res->negvol2; tsw->done = 1; tsw->skip = 1;
          tsw->err_step = tsw->nstep;
        }
      }
    }
    if (skip0 || res->negvol || res->negvol2 || !tsw->ale_on) return;
  } else {
    REMAP_GATE(B);
  }
  const int e = xcd_index();
  if (e >= B.nel) return;
  int nd[4];
  load_nd(B.ielnd, e, nd);
  const NbFc c = load_nbfc_packed(e, B.nbfc);
  double ax[4], ay[4], bx[4], by[4], own[4], nbv[4];
  remap_faces(B, nd, B.ts->dt, ax, ay, bx, by, own, nbv);
  B.dv[e] = sumflux_regs(e, 0, 1, c.nb, own, nbv);
  double f[4];
  fluxelvl_regs(e, 0, 1, c.nb, c.fc, B.cnwt, own, B.rho, f);
  st4(B.fcdm, e, f);
}

// R2: ale_kn_sumflux(fcdm) -> mass change, ale_kn_updatebasisel (volume and
// mass after the remap), ale_kn_fluxelvl(cnmass, fcdm, elenergy) -> face
// energy fluxes (ale_dr_advect.f90:190-215)
__global__ void __launch_bounds__(kBlock) k_r2(AleArgs B) {
  REMAP_GATE(B);
  const int e = xcd_index();
  if (e >= B.nel) return;
  const NbFc c = load_nbfc_packed(e, B.nbfc);
  const double totm = sumflux_el(e, 0, 1, c.nb, c.fc, B.fcdm);
  const double v0 = B.vol[e], m0 = B.mass[e];
  B.v0[e] = v0;
  B.m0[e] = m0;
  B.mass[e] = m0 + totm;
  double f[4];
  fluxelvl_el(e, 0, 1, c.nb, c.fc, B.cnmass, B.fcdm, B.en, f);
  st4(B.flux, e, f);
}

// R3: ale_kn_updateel (energy), the element part of the nodal remap
// (corner gathers, ale_kn_fluxbasisnd, corner mass update) and
// ale_dr_hydroupdate (geometry on the moved mesh, density, EOS)
__global__ void __launch_bounds__(kBlock) k_r3(AleArgs B) {
  REMAP_GATE(B);
  const int e = xcd_index();
  if (e >= B.nel) return;
  int nd[4];
  load_nd(B.ielnd, e, nd);
  const NbFc c = load_nbfc_packed(e, B.nbfc);
  // ale_kn_updateel
  const double s = sumflux_el(e, 0, 1, c.nb, c.fc, B.flux);
  const double m1 = B.mass[e], v0 = B.v0[e];
  double en = B.en[e];
  if (m1 > v0 * B.dencut) en = (en * B.m0[e] + s) / m1;
  B.en[e] = en;
  // corner mass before the remap (R4 reads the node velocities through
  // ielnd where the reference gathers them to corners)
  double cm[4];
  ld4(B.cnmass, e, cm);
  st4(B.cnm0, e, cm);
  // ale_kn_fluxbasisnd, corner part of ale_kn_massbasisnd
  double ax[4], ay[4], bx[4], by[4], own[4], nbv[4], fm[4];
  remap_faces(B, nd, B.ts->dt, ax, ay, bx, by, own, nbv);
  ld4(B.fcdm, e, fm);
  double cd[4], fl4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i1 = 0; i1 <= 1; ++i1) {
    const int i2 = i1 + 2;
    const int e1 = c.nb[i1], e2 = c.nb[i2];
    double w3 = B.fcdm[4 * e1 + c.fc[i1]], w4 = B.fcdm[4 * e2 + c.fc[i2]];
    if (e1 == e) w3 = 0.0;
    if (e2 == e) w4 = 0.0;
    const double w1 = w3 - fm[i1];
    const double w2 = w4 - fm[i2];
    w3 = 0.25 * (w1 - w2);
    cd[i1] = w3;
    cd[i2] = w3;
    w3 = 0.25 * (w1 + w2);
    fl4[0] = fl4[0] + w3; fl4[1] = fl4[1] + w3; fl4[2] = fl4[2] + w3; fl4[3] = fl4[3] + w3;
  }
  st4(B.cndm, e, cd);
  const int nel = B.nel;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    B.cnflux[j * nel + e] = fl4[j];
    B.cnm0T[j * nel + e] = cm[j];
  }
  const double cm1[4] = {cm[0] + fl4[0], cm[1] + fl4[1], cm[2] + fl4[2], cm[3] + fl4[3]};
  st4(B.cnmass, e, cm1);
  // ale_dr_hydroupdate on the moved mesh: density and EOS; the geometry
  // fields are rebuilt from the final nodes at the end of the batch (nothing
  // in a step reads them before F4 rewrites the ones the remap needs)
  Iso g;
  getiso(bx, by, g);
  if (g.vol < 0.0) atomicMax(&B.dtres->negvol, e + 1);
  const double r = m1 / g.vol;
  B.rho[e] = r;
  int mat = B.ielmat[e];
  mat = mat > 0 ? mat : 0;
  const int et = B.eos_itype[mat];
  const double* ep = B.eos_param + BL_NEOSPARAM * mat;
  B.pre[e] = eos_pressure(et, ep, r, en, B.pcut);
  B.cs2[e] = eos_cs2(et, ep, r, en, B.pcut, B.ccut);
}

// R4: ale_kn_fluxndvl for both velocity components (ale_dr_advect.f90:
// 241-268; the v flux reads the corner v gathered before the u update, and
// R5 updates the nodes after this kernel, so both read u and v at the nodes)
// With MIX (R1 does F5's work and the step has mixed cells), R4's first
// nbm8 workgroups are F4m, one thread per mixed cell, gated as F5 gated
// it: not skipped, whether or not this step remaps (ts->skip now includes
// the step's negative volumes, recorded by R1).  With fold7 they then do
// the mixed cells' part of the hydro update as F7 did, on the positions R5
// will give the nodes, x + dt w, computed here the same way.
template <bool MIX>
__global__ void __launch_bounds__(kBlock) k_r4(AleArgs B, FusedArgs A, FusedNodes O, int nbm8,
                                               int fold7) {
  if (MIX && (int)blockIdx.x < nbm8) {
    if (A.ts->skip) return;
    const int imx = (int)blockIdx.x * kBlock + threadIdx.x;
    if (imx >= A.mx.nmx) return;
    f4_mix_cell(A, O, A.ts, imx);
    if (fold7 && A.ts->ale_on && !A.dtres->negvol) {
      int nd[4];
      load_nd(A.ielnd, A.mx.imxel[imx] - 1, nd);
      const double dt = A.ts->dt;
      double px[4], py[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double wu = B.ubar[nd[c]], wv = B.vbar[nd[c]];
        if (B.zeul) { wu = -wu; wv = -wv; }
        px[c] = O.x[nd[c]] + dt * wu;
        py[c] = O.y[nd[c]] + dt * wv;
      }
      Iso g;
      getiso(px, py, g);
      mix_hydroupdate_iso(A.mx, imx, g, A.eos_itype, A.eos_param, A.pcut, A.ccut);
    }
    return;
  }
  REMAP_GATE(B);
  const int e = MIX ? xcd_block_of((int)blockIdx.x - nbm8, (int)gridDim.x - nbm8) * kBlock +
                          threadIdx.x
                    : xcd_index();
  if (e >= B.nel) return;
  const NbFc c = load_nbfc_packed(e, B.nbfc);
  double ou[4], ov[4];
  const int nel = B.nel;
  fluxndvl_uv(e, c, B.ielnd, B.cnm0, B.cndm, B.u, B.v, ou, ov);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    B.fu[j * nel + e] = ou[j];
    B.fv[j * nel + e] = ov[j];
  }
}

// R5: node sums (ale_kn_initbasisnd, ale_kn_calcbasisnd, ale_kn_massbasisnd
// in ALE order, ale_kn_cutbasisnd), ale_kn_activend with indstatus = 2
// (Eulerian), ale_kn_updatend for u and v (element order), and the moved
// node positions
// With FOLD7 (Eulerian, no per-step diagnostics), R5's first thread also
// does F7's check: a negative volume in the hydro update (R3's) records the
// error and ends the run at this step (ts->done; ts->skip is left for the
// next step's F1b, so that no workgroup of R5 depends on it).
template <bool FOLD7>
__global__ void __launch_bounds__(kBlock) k_r5(AleArgs B, DevTimestep* tsw) {
  if (FOLD7 && blockIdx.x == 0 && threadIdx.x == 0 && !tsw->skip && tsw->ale_on &&
      B.dtres->negvol) {
    tsw->err = 1; tsw->err_idx = -B.dtres->negvol; tsw->done = 1;
    tsw->err_step = tsw->nstep;
  }
  REMAP_GATE(B);
  const int n = xcd_index();
  if (n >= B.nnd) return;
  const int nel = B.nel;
  double nv0 = 0.0, nm0 = 0.0;
  const int b = B.ale_ptr[n], end = B.ale_ptr[n + 1];
  for (int k = b; k < end; ++k) {
    const int cc = B.ale_idx[k];
    nv0 = nv0 + 0.25 * B.v0[cc >> 2];
    nm0 = nm0 + B.cnm0T[(cc & 3) * nel + (cc >> 2)];
  }
  double nm1 = nm0;
  for (int k = b; k < end; ++k) {
    const int cc = B.ale_idx[k];
    nm1 = nm1 + B.cnflux[(cc & 3) * nel + (cc >> 2)];
  }
  const double cutm = B.dencut * nv0;
  double su = 0.0, sv = 0.0;
  for (int k = B.el_ptr[n]; k < B.el_ptr[n + 1]; ++k) {
    const int cc = B.el_idx[k];
    const int off = (cc & 3) * nel + (cc >> 2);
    su = su + B.fu[off];
    sv = sv + B.fv[off];
  }
  const int t = B.indtype[n];
  double u = B.u[n], v = B.v[n];
  if (t != -1 && t != -3 && nm1 > cutm) u = (u * nm0 + su) / nm1;
  if (t != -2 && t != -3 && nm1 > cutm) v = (v * nm0 + sv) / nm1;
  double wu = B.ubar[n], wv = B.vbar[n];
  if (B.zeul) { wu = -wu; wv = -wv; }
  const double dt = B.ts->dt;
  B.x[n] = B.x[n] + dt * wu;
  B.y[n] = B.y[n] + dt * wv;
  B.u[n] = u;
  B.v[n] = v;
}

// negative volume found by F2 (F3/F4 did not run) or F4 (F3 did): stop the
// batch at this step; the ALE kernels that follow are gated by skip.
// Workgroups from 1 on are F4m, one thread per mixed cell: they run when
// neither F2 nor F4 found a negative volume (the reference halts in the
// corrector's geometry before the component state), a test that does not
// depend on workgroup 0's update of ts.
__global__ void __launch_bounds__(kBlock) k_f5_end(FusedArgs A, FusedNodes O) {
  DevTimestep* ts = A.ts;
  const DtResult* res = A.dtres;
  if (blockIdx.x > 0) {
    if (ts->skip || res->negvol || res->negvol2) return;
    const int imx = ((int)blockIdx.x - 1) * kBlock + threadIdx.x;
    if (imx < A.mx.nmx) f4_mix_cell(A, O, ts, imx);
    return;
  }
  if (threadIdx.x != 0 || ts->skip) return;
  if (ens_pending(A, ts)) {  // the host reads the instance and code from epend
    ts->done = 1; ts->skip = 1;
    return;
  }
  if (res->negvol) {
    ts->err = 1; ts->err_idx = -res->negvol; ts->done = 1; ts->skip = 1;
    ts->err_step = ts->nstep;
    return;
  }
  ts->flips = ts->flips + 1;
  if (res->negvol2) {
    ts->err = 1; ts->err_idx = -res->negvol2; ts->done = 1; ts->skip = 1;
    ts->err_step = ts->nstep;
  }
}

// negative volume in the ALE hydro update (ale_dr_update.f90:44).
// Workgroups from 1 on are the mixed cells' part of the hydro update (one
// thread per cell, on the remapped mesh x, y); they run unless the update
// found a negative volume (the reference halts in its geometry before the
// component state), a test that does not depend on workgroup 0's update of
// ts.
__global__ void __launch_bounds__(kBlock) k_f7_aleend(FusedArgs A, const double* __restrict__ x,
                                                      const double* __restrict__ y) {
  DevTimestep* ts = A.ts;
  const DtResult* res = A.dtres;
  if (ts->skip || !ts->ale_on) return;
  if (blockIdx.x > 0) {
    if (res->negvol) return;
    const int imx = ((int)blockIdx.x - 1) * kBlock + threadIdx.x;
    if (imx < A.mx.nmx)
      mix_hydroupdate_cell(A.mx, imx, A.ielnd, x, y, A.eos_itype, A.eos_param, A.pcut, A.ccut);
    return;
  }
  if (threadIdx.x != 0) return;
  if (res->negvol) {
    ts->err = 1; ts->err_idx = -res->negvol; ts->done = 1; ts->skip = 1;
    ts->err_step = ts->nstep;
  }
}

}  // namespace bl

using namespace bl;

namespace {

FusedArgs make_args(bl_handle_t* h) {
  FusedArgs A{};
  const bl_config_t& c = h->cfg;
  A.nel = h->sz.nel;
  A.nnd = h->sz.nnd;
  bool dtnot = false, mid = false;
  for (int r = 0; r < h->nreg; ++r) {
    dtnot = dtnot || h->zdtnotreg[r];
    mid = mid || h->zmidlength[r];
  }
  A.reg_tables = (dtnot || mid) ? 1 : 0;
  A.zsp = c.zsp;
  A.zhg = c.zhg;
  A.reg_tables_force = (c.zsp || c.zhg) ? 1 : 0;
  A.zexist = c.zexist;
  A.write_geometry = c.zexist ? 1 : 0;
  A.zcut = c.zcut; A.zerocut = c.zerocut; A.dencut = c.dencut; A.accut2 = c.accut * c.accut;
  A.cvisc1 = c.cvisc1; A.cvisc2 = c.cvisc2; A.cfl_sf = c.cfl_sf; A.div_sf = c.div_sf;
  A.ccut = c.ccut; A.pcut = c.pcut;
  A.dt_g = c.dt_g; A.dt_min = c.dt_min; A.dt_max = c.dt_max; A.dt_initial = c.dt_initial;
  A.mintime = c.mintime; A.maxtime = c.maxtime;
  A.ielnd = h->I(BL_IELND); A.ielel = h->I(BL_IELEL);
  A.ielreg = h->I(BL_IELREG); A.ielmat = h->I(BL_IELMAT); A.indtype = h->I(BL_INDTYPE);
  A.hyd_ptr = h->d_hyd_ptr; A.hyd_idx = h->d_hyd_idx; A.outer = h->d_outer;
  A.zdtnotreg = h->d_zdtnotreg; A.zmidlength = h->d_zmidlength;
  A.kappareg = h->d_kappareg; A.pmeritreg = h->d_pmeritreg;
  A.eos_itype = h->d_eos_itype; A.eos_param = h->d_eos_param;
  A.rho = h->R(BL_ELDENSITY); A.en = h->R(BL_ELENERGY); A.pre = h->R(BL_ELPRESSURE);
  A.cs2 = h->R(BL_ELCS2); A.vol = h->R(BL_ELVOLUME); A.mass = h->R(BL_ELMASS);
  A.rho_o = A.rho; A.en_o = A.en; A.pre_o = A.pre; A.cs2_o = A.cs2;
  A.cnmass = h->R(BL_CNMASS); A.spmass = h->R(BL_SPMASS); A.cnwt = h->R(BL_CNWT);
  A.cnx = h->R(BL_CNX); A.cny = h->R(BL_CNY);
  A.a1 = h->R(BL_A1); A.a2 = h->R(BL_A2); A.a3 = h->R(BL_A3);
  A.b1 = h->R(BL_B1); A.b2 = h->R(BL_B2); A.b3 = h->R(BL_B3);
  A.cnviscx = h->R(BL_CNVISCX); A.cnviscy = h->R(BL_CNVISCY); A.elvisc = h->R(BL_ELVISC);
  A.cnfx = h->R(BL_RSCRATCH23); A.cnfy = h->R(BL_RSCRATCH24);
  A.fxT = h->cterm[0]; A.fyT = h->cterm[1]; A.wtT = h->cterm[2]; A.mT = h->cterm[3];
  A.ubar = h->R(BL_RSCRATCH14); A.vbar = h->R(BL_RSCRATCH15);
  A.part = h->d_part; A.parti = h->d_parti; A.dtres = h->d_dtres;
  A.work = h->d_work;
  A.rest = h->d_rest; A.vstats = h->d_vstats;
  A.fold = 0; A.rest_prev = nullptr; A.rest_next = nullptr;
  A.ts = h->d_ts;
  A.hist = h->d_hist;
  A.hist_cap = (int)h->hist_cap;
  A.mx = make_mix(h);
  return A;
}

AleArgs make_ale(bl_handle_t* h, const FusedNodesOut& O) {
  AleArgs B{};
  B.nel = h->sz.nel;
  B.nnd = h->sz.nnd;
  B.zeul = h->cfg.zeul;
  B.zerocut = h->cfg.zerocut; B.dencut = h->cfg.dencut;
  B.pcut = h->cfg.pcut; B.ccut = h->cfg.ccut;
  B.ielnd = h->I(BL_IELND); B.ielel = h->I(BL_IELEL); B.ielfc = h->I(BL_IELFC);
  B.nbfc = h->d_nbfc;
  B.indtype = h->I(BL_INDTYPE); B.ielmat = h->I(BL_IELMAT);
  B.eos_itype = h->d_eos_itype; B.eos_param = h->d_eos_param;
  B.ale_ptr = h->d_ale_ptr; B.ale_idx = h->d_ale_idx;
  B.el_ptr = h->d_el_ptr; B.el_idx = h->d_el_idx;
  B.x = O.x; B.y = O.y; B.u = O.u; B.v = O.v;
  B.ubar = h->R(BL_RSCRATCH14); B.vbar = h->R(BL_RSCRATCH15);
  B.cnwt = h->R(BL_CNWT); B.rho = h->R(BL_ELDENSITY); B.vol = h->R(BL_ELVOLUME);
  B.mass = h->R(BL_ELMASS); B.en = h->R(BL_ELENERGY); B.pre = h->R(BL_ELPRESSURE);
  B.cs2 = h->R(BL_ELCS2); B.cnmass = h->R(BL_CNMASS);
  B.cnx = h->R(BL_CNX); B.cny = h->R(BL_CNY);
  B.a1 = h->R(BL_A1); B.a2 = h->R(BL_A2); B.a3 = h->R(BL_A3);
  B.b1 = h->R(BL_B1); B.b2 = h->R(BL_B2); B.b3 = h->R(BL_B3);
  double** t = h->ale_tmp;
  B.fcdm = t[0]; B.flux = t[1]; B.cnm0 = t[2]; B.cndm = t[3];
  B.cnflux = t[4]; B.fu = t[5]; B.fv = t[6];
  B.dv = t[7]; B.v0 = t[8]; B.m0 = t[9]; B.cnm0T = t[10];
  B.dtres = h->d_dtres;
  B.ts = h->d_ts;
  return B;
}

// F1 folds into F4 in a Lagrangian step of pure cells: the next step's
// start state is then complete when F4 ends (no remap and no mixed-cell
// update follows)
bool fold_f1(const bl_handle_t* h) {
  return !h->cfg.zexist && h->sz.ncp == 0 && h->ens_n == 0;
}

// no F5 in a folded Lagrangian step without per-step diagnostics (whose
// records need the step's skip flag before the next step): F1b records the
// step before's end, the host the batch's last
bool end_in_f1b(const bl_handle_t* h) {
  return fold_f1(h) && !h->diag_on;
}

// the tile step runs a folded Lagrangian step when the mesh has a tiling
// (bl_tiles.hip) and tiling is on; with per-step diagnostics F5 ends each
// step and the records follow it
bool use_tiles(const bl_handle_t* h) {
  return fold_f1(h) && h->tiles_on && h->tile_state == 1;
}

// an ensemble's step on the tiling (tiles within instances): F5 still ends
// each step, and E1's terms go to the tiles' slots and the worklists
bool ens_tiles(const bl_handle_t* h) {
  return h->ens_n > 0 && h->tiles_on && h->tile_state == 1;
}

// the tile step's arguments for a step from node and element set p: it
// reads element set p and writes set 1 - p
void tile_args(bl_handle_t* h, FusedArgs& A, int p) {
  const int nel = h->sz.nel;
  A.ntile = h->ntile;
  A.nseg = h->ntile;
  A.tile_eptr = h->d_tile_eptr;
  A.tile_nptr = h->d_tile_nptr;
  A.tile_bounds = h->d_tile_bounds;
  A.tile_nodes = h->d_tile_nodes;
  A.tile_ncm = h->d_tile_ncm;
  A.tile_el = h->d_tile_el;
  A.tile_nd = h->d_tile_nd;
  A.rho = h->elset[p][0]; A.en = h->elset[p][1]; A.pre = h->elset[p][2]; A.cs2 = h->elset[p][3];
  A.rho_o = h->elset[1 - p][0]; A.en_o = h->elset[1 - p][1];
  A.pre_o = h->elset[1 - p][2]; A.cs2_o = h->elset[1 - p][3];
  A.rest8 = h->d_rest8 + (size_t)p * nel;
  A.rest8_prev = h->d_rest8 + (size_t)(1 - p) * nel;
  A.rest8_next = h->d_rest8 + (size_t)(1 - p) * nel;
  A.fold = 1;
  A.end_in_f1b = 1;
}

// the two at-rest masks of a folded run: m[p] is the mask of a step that
// starts from node set p
void rest_masks(bl_handle_t* h, uint64_t* m[2]) {
  m[0] = h->d_rest;
  m[1] = h->d_rest + 4 * (size_t)nblocks(std::max(h->sz.nel, 1));
}

// the first F1 of a folded run call, on node set p, outside the step graphs
int launch_f1_prologue(bl_handle_t* h, int p, int stop_at_time_end) {
  FusedArgs A = make_args(h);
  A.stop_at_time_end = stop_at_time_end;
  A.time_end = h->cfg.time_end;
  if (use_tiles(h)) {
    tile_args(h, A, p);
    A.rest8_next = h->d_rest8 + (size_t)p * h->sz.nel;  // the first step's own bytes
    FusedNodes N{h->nodeset[p][0], h->nodeset[p][1], h->nodeset[p][2], h->nodeset[p][3]};
    BL_LAUNCH(h, "f1_seg", k_f1_seg, (unsigned)h->ntile, kBlock, A, N, EnsArgs{});
    return BL_SUCCESS;
  }
  uint64_t* m[2];
  rest_masks(h, m);
  A.fold = 1;
  A.rest = m[p];
  FusedNodes N{h->nodeset[p][0], h->nodeset[p][1], h->nodeset[p][2], h->nodeset[p][3]};
  BL_LAUNCH(h, "f1_getdt", k_f1_getdt, nblocks(h->sz.nel), kBlock, A, N);
  return BL_SUCCESS;
}

// one step (parity p: node set p -> node set 1-p), launched on h->stream
int launch_step(bl_handle_t* h, int p, int stop_at_time_end) {
  FusedArgs A = make_args(h);
  A.stop_at_time_end = stop_at_time_end;
  A.time_end = h->cfg.time_end;
  const int nel = h->sz.nel, nnd = h->sz.nnd;
  const unsigned nbe = nblocks(nel), nbn = nblocks(nnd);
  FusedNodes N{h->nodeset[p][0], h->nodeset[p][1], h->nodeset[p][2], h->nodeset[p][3]};
  FusedNodesOut O{h->nodeset[1 - p][0], h->nodeset[1 - p][1], h->nodeset[1 - p][2],
                  h->nodeset[1 - p][3]};
  FusedNodes Oc{O.x, O.y, O.u, O.v};
  if (use_tiles(h)) {
    tile_args(h, A, p);
    A.end_in_f1b = end_in_f1b(h) ? 1 : 0;  // with per-step diagnostics F5 ends the step
    const unsigned nseg = (unsigned)A.nseg;
    const unsigned nbl = std::min(nseg, (unsigned)h->tune_lim_grid);
    BL_LAUNCH(h, "f1_limiter", k_f1_limiter, nbl, (unsigned)h->tune_lim_block, A, N, (int)nseg);
    BL_LAUNCH(h, "f1b_select", k_f1b_select, 1, kF1bThreads, A, (int)nseg, (int)nbl);
    BL_LAUNCH(h, "lag_tile", k_lag_tile, (unsigned)h->ntile, kBlock, A, N, O, EnsArgs{});
    if (end_in_f1b(h)) return BL_SUCCESS;
    // per-step diagnostics: F5 records the step's end, then the record of
    // the node and element sets the step wrote
    BL_LAUNCH(h, "f5_end", k_f5_end, 1, kBlock, A, Oc);
    return launch_diag(h, O.x, O.y, O.u, O.v, A.rho_o, A.en_o, A.pre_o);
  }
  const unsigned nbl = std::min(nbe, (unsigned)h->tune_lim_grid);
  A.end_in_f1b = end_in_f1b(h) ? 1 : 0;
  if (fold_f1(h)) {
    // F1 ran in the previous step's F4 (or in the run call's prologue)
    uint64_t* m[2];
    rest_masks(h, m);
    A.fold = 1;
    A.rest = m[p];
    A.rest_prev = m[1 - p];
    A.rest_next = m[1 - p];
  } else {
    BL_LAUNCH(h, "f1_getdt", k_f1_getdt, nbe, kBlock, A, N);
  }
  BL_LAUNCH(h, "f1_limiter", k_f1_limiter, nbl, (unsigned)h->tune_lim_block, A, N, (int)nbe);
  BL_LAUNCH(h, "f1b_select", k_f1b_select, 1, kF1bThreads, A, (int)nbe, (int)nbl);
  // F2 and F5 (or R1) take a workgroup per 256 mixed cells for the
  // component work; F2 puts them first, rounded up to whole XCD rounds
  const unsigned nbm = A.mx.ncp > 0 ? nblocks(A.mx.nmx) : 0;
  const unsigned nbm8 = (nbm + kNumXcd - 1) / kNumXcd * kNumXcd;
  BL_LAUNCH(h, "f2_force", k_f2_force, nbm8 + nbe, kBlock, A, N, (int)nbm8);
  BL_LAUNCH(h, "f3_accel", k_f3_accel, nbn, kBlock, A, N, O);
  BL_LAUNCH(h, "f4_state", k_f4_state, nbe, kBlock, A, Oc);
  // Eulerian ALE: F5 runs inside R1 and F4m inside R4
  const bool fold_f5 = h->cfg.zexist && h->cfg.zeul;
  if (!fold_f5 && !A.end_in_f1b) BL_LAUNCH(h, "f5_end", k_f5_end, 1 + nbm, kBlock, A, Oc);
  if (h->cfg.zexist) {
    // the remap on the node set this step wrote
    AleArgs B = make_ale(h, O);
    // node relaxation: its mesh velocity overwrites ubar (B.zeul = 0)
    if (!h->cfg.zeul &&
        mesh_motion(h, O.x, O.y, h->R(BL_RSCRATCH14), h->R(BL_RSCRATCH15), 0.0, h->d_ts))
      return BL_FAILURE;
    if (fold_f5)
      BL_LAUNCH(h, "r1_volflux", k_r1<true>, nbe, kBlock, B, h->d_ts);
    else
      BL_LAUNCH(h, "r1_volflux", k_r1<false>, nbe, kBlock, B, h->d_ts);
    BL_LAUNCH(h, "r2_massflux", k_r2, nbe, kBlock, B);
    BL_LAUNCH(h, "r3_energy_update", k_r3, nbe, kBlock, B);
    // with no per-step diagnostics, F7 folds into R4 (mixed cells) and R5
    const bool fold_f7 = fold_f5 && !h->diag_on;
    if (fold_f5 && nbm > 0)
      BL_LAUNCH(h, "r4_momflux", k_r4<true>, nbm8 + nbe, kBlock, B, A, Oc, (int)nbm8,
                fold_f7 ? 1 : 0);
    else
      BL_LAUNCH(h, "r4_momflux", k_r4<false>, nbe, kBlock, B, A, Oc, 0, 0);
    if (fold_f7) {
      BL_LAUNCH(h, "r5_nodes", k_r5<true>, nbn, kBlock, B, h->d_ts);
    } else {
      BL_LAUNCH(h, "r5_nodes", k_r5<false>, nbn, kBlock, B, h->d_ts);
      // F7, with the mixed cells' hydro update in its appended workgroups
      BL_LAUNCH(h, "f7_aleend", k_f7_aleend, 1 + nbm, kBlock, A, (const double*)O.x,
                (const double*)O.y);
    }
  }
  return launch_diag(h, O.x, O.y, O.u, O.v);
}

int ensure_nodesets(bl_handle_t* h) {
  if (!h->d_work) {
    // an ensemble's per-instance worklist segments
    BL_HIP(hipMalloc(&h->d_work, sizeof(int32_t) * std::max(h->sz.nel, 1)));
  }
  if (!h->d_rest) {
    // one word per wave of F1's grid, the partial last workgroup included,
    // twice (the current and the previous step's with F1 folded into F4)
    const size_t words = 8 * (size_t)nblocks(std::max(h->sz.nel, 1));
    BL_HIP(hipMalloc(&h->d_rest, sizeof(uint64_t) * words));
    BL_HIP(hipMalloc(&h->d_vstats, sizeof(unsigned long long) * kStatWords));
    BL_HIP(hipMemset(h->d_rest, 0, sizeof(uint64_t) * words));
    BL_HIP(hipMemset(h->d_vstats, 0, sizeof(unsigned long long) * kStatWords));
  }
  if (!h->cterm[0])
    for (int k = 0; k < 4; ++k)
      BL_HIP(hipMalloc(&h->cterm[k], sizeof(double) * 4 * std::max(h->sz.nel, 1)));
  if (h->cfg.zexist && !h->ale_tmp[0])
    for (int k = 0; k < 11; ++k)
      BL_HIP(hipMalloc(&h->ale_tmp[k], sizeof(double) * ((k < 7 || k == 10) ? 4 : 1) *
                                           std::max(h->sz.nel, 1)));
  if (h->nodeset[1][0]) return BL_SUCCESS;
  const int ids[4] = {BL_NDX, BL_NDY, BL_NDU, BL_NDV};
  for (int k = 0; k < 4; ++k) {
    h->nodeset[0][k] = h->R(ids[k]);
    BL_HIP(hipMalloc(&h->nodeset[1][k], sizeof(double) * std::max(h->sz.nnd2, 1)));
  }
  h->cur_set = 0;
  return BL_SUCCESS;
}

// Forget which elements' viscosity fields hold F1's zeros: at the start of
// every run call, since between calls the host may have written them (an
// upload, a per-driver call, a write through bl_field_device_ptr).  The
// first step of the call writes every element's.  Outside any capture.
// Everything a run call starts from, in one launch: the at-rest masks (bits
// and bytes), the counters, the negative-volume flags the last call's last
// step may have left (F1L and F1b read them as the step before's outcome),
// an ensemble's worklist lengths (a folded run call ends with the next
// step's lists) and, with set_ts, the device timestep.
__global__ void k_call_start(DevTimestep init, int set_ts, DevTimestep* __restrict__ ts,
                             DtResult* __restrict__ dtres, uint64_t* __restrict__ rest,
                             int64_t nrest, unsigned long long* __restrict__ vstats,
                             uint8_t* __restrict__ rest8, int64_t nrest8,
                             uint32_t* __restrict__ cnt, int ncnt) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (i0 == 0) {
    *dtres = DtResult{};
    if (set_ts) *ts = init;
  }
  for (int64_t i = i0; i < nrest; i += stride) rest[i] = 0;
  for (int64_t i = i0; i < kStatWords; i += stride) vstats[i] = 0;
  for (int64_t i = i0; i < ncnt; i += stride) cnt[i] = 0;
  // the bytes: whole 8-byte words (hipMalloc'd, so aligned), then the tail
  uint64_t* r8 = reinterpret_cast<uint64_t*>(rest8);
  for (int64_t i = i0; i < nrest8 / 8; i += stride) r8[i] = 0;
  for (int64_t i = (nrest8 / 8) * 8 + i0; i < nrest8; i += stride) rest8[i] = 0;
}

// Forget which elements' viscosity fields hold F1's zeros (and the rest of
// what k_call_start clears); with init, also write the device timestep.
// keep_masks: a run call that continues the last one keeps the masks
int clear_rest(bl_handle_t* h, const DevTimestep* init = nullptr, bool keep_masks = false) {
  const int64_t nrest = keep_masks ? 0 : 8 * (int64_t)nblocks(std::max(h->sz.nel, 1));
  const int64_t nrest8 = (h->d_rest8 && !keep_masks) ? 2 * (int64_t)std::max(h->sz.nel, 1) : 0;
  const int ncnt = (h->ens_n > 0 && h->d_cnt) ? h->ens_n : 0;
  const int64_t most = std::max(std::max(nrest, nrest8 / 8), (int64_t)std::max(kStatWords, ncnt));
  const unsigned grid = (unsigned)std::min<int64_t>(1024, (most + kBlock - 1) / kBlock);
  DevTimestep none{};
  BL_LAUNCH(h, "call_start", k_call_start, std::max(grid, 1u), kBlock, init ? *init : none,
            init ? 1 : 0, h->d_ts, h->d_dtres, h->d_rest, nrest, h->d_vstats, h->d_rest8,
            nrest8, h->d_cnt, ncnt);
  return BL_SUCCESS;
}

void point_fields_at(bl_handle_t* h, int s) {
  h->f[BL_NDX] = h->nodeset[s][0];
  h->f[BL_NDY] = h->nodeset[s][1];
  h->f[BL_NDU] = h->nodeset[s][2];
  h->f[BL_NDV] = h->nodeset[s][3];
  h->cur_set = s;
}

// the element state after a tile-step batch: set s, like the nodes
void point_elements_at(bl_handle_t* h, int s) {
  h->f[BL_ELDENSITY] = h->elset[s][0];
  h->f[BL_ELENERGY] = h->elset[s][1];
  h->f[BL_ELPRESSURE] = h->elset[s][2];
  h->f[BL_ELCS2] = h->elset[s][3];
  h->el_cur = s;
}

// Before a tile-step batch: the tiling (built when the mesh's connectivity
// changed), and element set cur_set holding the registry's element state
// (a batch of another step kind moved the node sets but kept the elements
// in place).  Outside any capture.
int prepare_tiles(bl_handle_t* h) {
  if (!h->tiles_on || !(fold_f1(h) || h->ens_n > 0)) return BL_SUCCESS;
  if (h->tile_state == 0) {
    if (build_tiles(h)) return BL_FAILURE;
    invalidate_graph(h);
  }
  if (h->tile_state != 1) return BL_SUCCESS;
  if (h->el_cur != h->cur_set) {
    for (int k = 0; k < 4; ++k) std::swap(h->elset[0][k], h->elset[1][k]);
    h->el_cur = h->cur_set;
    invalidate_graph(h);
  }
  return BL_SUCCESS;
}

// the packed neighbour/face table of the fused remap, from ielel and ielfc
// (rebuilt at a capture after they were uploaded or the maps rebuilt)
constexpr int kMinHistory = 8192;  // dt-history records a graph run can hold

int build_nbfc(bl_handle_t* h) {
  const int nel = h->sz.nel;
  std::vector<int32_t> el(4 * (size_t)nel), fc(4 * (size_t)nel), pk(4 * (size_t)nel);
  if (bl_download(h, BL_IELEL, el.data(), 4 * (int64_t)nel) ||
      bl_download(h, BL_IELFC, fc.data(), 4 * (int64_t)nel))
    return BL_FAILURE;
  for (size_t k = 0; k < pk.size(); ++k) {
    if (el[k] < 1 || el[k] > nel || fc[k] < 1 || fc[k] > 4) {
      set_error("ielel/ielfc out of range for the remap (element " +
                std::to_string(k / 4 + 1) + ")");
      return BL_FAILURE;
    }
    pk[k] = ((el[k] - 1) << 2) | (fc[k] - 1);
  }
  if (!h->d_nbfc)
    BL_HIP(hipMalloc(&h->d_nbfc, sizeof(int32_t) * std::max<size_t>(pk.size(), 4)));
  BL_HIP(hipMemcpy(h->d_nbfc, pk.data(), sizeof(int32_t) * pk.size(), hipMemcpyHostToDevice));
  h->nbfc_stale = false;
  return BL_SUCCESS;
}

// Step graphs: consecutive graph launches leave a gap of several
// microseconds on the device between them, kernels inside one graph none,
// so a run call launches the fewest graphs its steps need (8, 4, 2 or 1
// steps each, largest first; 8 steps took the Sedov step from 0.1524 to
// 0.1486 ms), the last one with the batch's end in it.  Graphs of 16 and
// 32 steps made a 20-step call slower (0.127 against 0.126 ms a step).
constexpr int kGraphLogs = 4;
constexpr int kGraphs = 4 * kGraphLogs;
static_assert(std::extent<decltype(bl::Handle::graph_exec)>::value >= (size_t)kGraphs,
              "the handle's graph table holds every step graph");
inline int graph_index(int fin, int si, int p) { return 2 * kGraphLogs * fin + 2 * si + p; }

int launch_batch_end(bl_handle_t* h, const DevTimestep* d_ts, int cur);

// capture graph_exec[graph_index(fin, si, p)]; before any launch of a call
int capture_graph(bl_handle_t* h, int fin, int si, int p, int stop_at_time_end) {
  const bool timing = h->timing;
  h->timing = false;  // no event records inside a capture
  hipGraph_t g = nullptr;
  BL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = BL_SUCCESS;
  for (int k = 0; k < (1 << si) && rc == BL_SUCCESS; ++k)
    rc = launch_step(h, p ^ (k & 1), stop_at_time_end);
  if (rc == BL_SUCCESS && fin) rc = launch_batch_end(h, h->d_ts, 0);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->timing = timing;
  if (rc != BL_SUCCESS || e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    set_error("step graph capture failed");
    return BL_FAILURE;
  }
  e = hipGraphInstantiate(&h->graph_exec[graph_index(fin, si, p)], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return BL_FAILURE;
  }
  return BL_SUCCESS;
}

// the graphs of a run call of nsteps from node set p, in launch order:
// (fin, si, parity) of each
std::vector<int> graph_plan(int nsteps, int p) {
  std::vector<int> plan;
  for (int rem = nsteps; rem > 0;) {
    int si = kGraphLogs - 1;
    while ((1 << si) > rem) --si;
    const int fin = rem == (1 << si) ? 1 : 0;
    plan.push_back(graph_index(fin, si, p));
    p ^= (1 << si) & 1;
    rem -= 1 << si;
  }
  return plan;
}

// The end of a batch whose last step left its end open (no F1b after it),
// on the device, so that the host reads the timestep once after all the
// call's work
__global__ void k_batch_end(DevTimestep* ts, const DtResult* res) {
  if (threadIdx.x == 0) step_end(ts, res);
}

// The registry's geometry fields after a Lagrangian batch: F4 keeps them in
// registers, so they are rebuilt once from the final coordinates; the last
// step's corrector corner forces go back to the element-major rscratch23/24
// from the corner-major copy fxT, fyT (fxT null: the tile step's last step
// wrote them there itself).  The final node set is (ts->set0 + ts->flips) & 1
// of (x0, y0), (x1, y1): read on the device, after the batch; with ts null
// it is set cur.  (Writing the geometry in the batch's last tile step instead
// slowed every step's tile kernels by more than this pass costs once.)
__global__ void k_materialize(int nel, const int32_t* __restrict__ ielnd,
                              const double* __restrict__ x0, const double* __restrict__ y0,
                              const double* __restrict__ x1, const double* __restrict__ y1,
                              const DevTimestep* __restrict__ ts, int cur,
                              double* __restrict__ cnx, double* __restrict__ cny,
                              double* __restrict__ a1, double* __restrict__ a2,
                              double* __restrict__ a3, double* __restrict__ b1,
                              double* __restrict__ b2, double* __restrict__ b3,
                              double* __restrict__ cnwt, double* __restrict__ vol,
                              const double* __restrict__ fxT, const double* __restrict__ fyT,
                              double* __restrict__ cnfx, double* __restrict__ cnfy) {
  int e = xcd_index();
  if (e >= nel) return;
  const bool other = ((ts ? ts->set0 + ts->flips : cur) & 1) != 0;
  const double* __restrict__ ndx = other ? x1 : x0;
  const double* __restrict__ ndy = other ? y1 : y0;
  if (fxT) {
    double fx[4], fy[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      fx[j] = fxT[j * nel + e];
      fy[j] = fyT[j * nel + e];
    }
    st4_nt(cnfx, e, fx);
    st4_nt(cnfy, e, fy);
  }
  int nd[4];
  load_nd(ielnd, e, nd);
  double x[4], y[4];
  load4(ndx, nd, x);
  load4(ndy, nd, y);
  Iso g;
  getiso(x, y, g);
  // (streaming stores: the next call's first step finds its state cached)
  st4_nt(cnx, e, x);
  st4_nt(cny, e, y);
  st_nt(a1, e, g.a1); st_nt(a2, e, g.a2); st_nt(a3, e, g.a3);
  st_nt(b1, e, g.b1); st_nt(b2, e, g.b2); st_nt(b3, e, g.b3);
  st4_nt(cnwt, e, g.wt);
  st_nt(vol, e, g.vol);
}

// The call's results for the host, in pinned coherent host memory, so that
// no copy follows the batch: the timestep and the dt history (3 values per
// step that selected a dt)
__global__ void k_publish(const DevTimestep* __restrict__ ts, const double* __restrict__ hist,
                          DevTimestep* __restrict__ out_ts, double* __restrict__ out_hist,
                          int cap) {
  const int n = min(ts->hist_n, cap);
  for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) out_hist[i] = hist[i];
  if (threadIdx.x == 0) *out_ts = *ts;
}

// The batch's end: the open end of its last step, then the registry's
// geometry at the final node set (from ts after the steps, or set cur with
// ts null; the tile step's last step wrote rscratch23/24 itself), then the
// results for the host
int launch_batch_end(bl_handle_t* h, const DevTimestep* d_ts, int cur) {
  if (end_in_f1b(h))
    BL_LAUNCH(h, "batch_end", k_batch_end, 1, 64, h->d_ts, (const DtResult*)h->d_dtres);
  const int nel = h->sz.nel;
  const bool tiles = use_tiles(h);
  BL_LAUNCH(h, "materialize", k_materialize, nblocks(nel), kBlock, nel, h->I(BL_IELND),
            h->nodeset[0][0], h->nodeset[0][1], h->nodeset[1][0], h->nodeset[1][1], d_ts, cur,
            h->R(BL_CNX), h->R(BL_CNY), h->R(BL_A1), h->R(BL_A2), h->R(BL_A3), h->R(BL_B1),
            h->R(BL_B2), h->R(BL_B3), h->R(BL_CNWT), h->R(BL_ELVOLUME),
            tiles ? nullptr : (const double*)h->cterm[0],
            tiles ? nullptr : (const double*)h->cterm[1], h->R(BL_RSCRATCH23),
            h->R(BL_RSCRATCH24));
  BL_LAUNCH(h, "publish", k_publish, 1, kBlock, (const DevTimestep*)h->d_ts,
            (const double*)h->d_hist, h->h_pub_ts, h->h_pub_hist, (int)h->hist_cap);
  return BL_SUCCESS;
}

void fill_err(bl_error_t* err, int ierr, int iout, const char* msg) {
  if (!err) return;
  err->ierr = ierr;
  err->iout = iout;
  std::memset(err->serr, ' ', BL_SLEN);
  size_t n = std::strlen(msg);
  std::memcpy(err->serr, msg, n < (size_t)BL_SLEN ? n : (size_t)BL_SLEN);
}

}  // namespace

namespace {

EnsArgs make_ens(bl_handle_t* h) {
  EnsArgs E{};
  E.ninst = h->ens_n;
  E.eoff = h->d_eoff;
  E.inst = static_cast<InstTs*>(h->d_inst);
  E.inst_dt = h->d_inst_dt;
  E.inst_rdt = h->d_inst_dt + h->ens_n;
  E.cnt = h->d_cnt;
  const size_t nb = nblocks(h->sz.nel);
  E.part = h->d_epart;
  E.parti = h->d_eparti;
  E.lboff = h->d_lboff;
  E.lpart = h->d_epart + 4 * nb;
  E.lparti = h->d_eparti + 4 * nb;
  E.hist = h->d_ehist;
  E.hist_cap = (int)h->ehist_cap;
  E.toff = ens_tiles(h) ? h->d_ens_toff : nullptr;
  return E;
}

// one ensemble step (parity p: node set p -> node set 1-p)
int launch_ensemble_step(bl_handle_t* h, int p) {
  FusedArgs A = make_args(h);
  A.stop_at_time_end = 0;
  A.time_end = h->cfg.time_end;
  A.einst = h->d_einst;
  A.ninst = h->d_ninst;
  A.inst_dt = h->d_inst_dt;
  A.inst_rdt = h->d_inst_dt + h->ens_n;
  EnsArgs E = make_ens(h);
  const int nel = h->sz.nel, nnd = h->sz.nnd;
  const unsigned nbe = nblocks(nel), nbn = nblocks(nnd);
  FusedNodes N{h->nodeset[p][0], h->nodeset[p][1], h->nodeset[p][2], h->nodeset[p][3]};
  FusedNodesOut O{h->nodeset[1 - p][0], h->nodeset[1 - p][1], h->nodeset[1 - p][2],
                  h->nodeset[1 - p][3]};
  FusedNodes Oc{O.x, O.y, O.u, O.v};
  if (ens_tiles(h)) {
    // E1 ran in the previous step's tile step (or in the run call's
    // prologue); the tile step reads element set p and writes 1 - p
    tile_args(h, A, p);
    A.end_in_f1b = 0;  // F5 ends each step
    BL_LAUNCH(h, "e1_limiter", k_e1_limiter, (unsigned)h->ens_lblocks,
              (unsigned)h->tune_lim_block, A, N, E);
    BL_LAUNCH(h, "e1b_select", k_e1b_select, (unsigned)h->ens_n, kBlock, A, E, (int)h->ntile);
    BL_LAUNCH(h, "lag_tile", k_lag_tile, (unsigned)h->ntile, kBlock, A, N, O, E);
    BL_LAUNCH(h, "f5_end", k_f5_end, 1, kBlock, A, Oc);
    return BL_SUCCESS;
  }
  // E1 ran in the previous step's F4 (or in the run call's prologue)
  uint64_t* m[2];
  rest_masks(h, m);
  A.fold = 1;
  A.rest = m[p];
  A.rest_prev = m[1 - p];
  A.rest_next = m[1 - p];
  BL_LAUNCH(h, "e1_limiter", k_e1_limiter, (unsigned)h->ens_lblocks, (unsigned)h->tune_lim_block,
            A, N, E);
  BL_LAUNCH(h, "e1b_select", k_e1b_select, (unsigned)h->ens_n, kBlock, A, E, (int)nbe);
  BL_LAUNCH(h, "f2_force", k_f2_force, nbe, kBlock, A, N, 0);
  BL_LAUNCH(h, "f3_accel", k_f3_accel, nbn, kBlock, A, N, O);
  BL_LAUNCH(h, "f4_state", k_f4_state_ens, nbe, kBlock, A, Oc, E);
  BL_LAUNCH(h, "f5_end", k_f5_end, 1, kBlock, A, Oc);
  return BL_SUCCESS;
}

int capture_ensemble(bl_handle_t* h) {
  invalidate_graph(h);
  bool timing = h->timing;
  h->timing = false;
  for (int p = 0; p < 2; ++p) {
    hipGraph_t g = nullptr;
    BL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_ensemble_step(h, p);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc != BL_SUCCESS || e != hipSuccess) {
      if (g) hipGraphDestroy(g);
      h->timing = timing;
      set_error("ensemble step graph capture failed");
      return BL_FAILURE;
    }
    e = hipGraphInstantiate(&h->graph_exec[p], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      h->timing = timing;
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      return BL_FAILURE;
    }
  }
  h->timing = timing;
  h->graph_stop_at_end = -2;  // marks the ensemble graphs
  return BL_SUCCESS;
}

}  // namespace

extern "C" {

int bl_ensemble_run(bl_handle_t* h, int32_t nsteps, double* dt_hist, int32_t* idt_hist,
                    int32_t* ctl_hist, int32_t* steps_taken, bl_error_t* err) {
  bl_touch(h);
  if (!h || h->ens_n <= 0) {
    set_error("bl_ensemble_run: not an ensemble handle");
    return BL_FAILURE;
  }
  fill_err(err, BL_SUCCESS, 0, "");
  if (nsteps <= 0) { if (steps_taken) *steps_taken = 0; return BL_SUCCESS; }
  if (ensure_nodesets(h) || prepare_tiles(h) || clear_rest(h)) return BL_FAILURE;
  if (h->ehist_cap < nsteps) {
    // as bl_run_graph: room for runs of up to 1024 steps without re-capture
    const int cap = std::max(nsteps, 1024);
    if (h->d_ehist) hipFree(h->d_ehist);
    h->d_ehist = nullptr;
    BL_HIP(hipMalloc(&h->d_ehist, sizeof(double) * 3 * (size_t)cap * h->ens_n));
    h->ehist_cap = cap;
    invalidate_graph(h);
  }
  if (!h->graph_exec[0] || h->graph_stop_at_end != -2)
    if (capture_ensemble(h)) return BL_FAILURE;
  DevTimestep* d = h->h_ts;
  BL_HIP(hipMemcpyAsync(d, h->d_ts, sizeof(DevTimestep), hipMemcpyDeviceToHost, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  const int start = d->nstep;
  d->done = 0; d->skip = 0; d->err = 0; d->err_step = 0; d->err_idx = 0; d->flips = 0;
  d->hist_base = start;
  d->epend = kNoPending;
  d->last_step = start + nsteps;
  BL_HIP(hipMemcpyAsync(h->d_ts, d, sizeof(DevTimestep), hipMemcpyHostToDevice, h->stream));
  int p = h->cur_set;
  {
    // E1 of the first step; every later step's runs in the F4 before it
    FusedArgs A = make_args(h);
    A.stop_at_time_end = 0;
    A.time_end = h->cfg.time_end;
    A.einst = h->d_einst;
    A.ninst = h->d_ninst;
    A.inst_dt = h->d_inst_dt;
    A.inst_rdt = h->d_inst_dt + h->ens_n;
    uint64_t* m[2];
    rest_masks(h, m);
    A.fold = 1;
    A.rest = m[p];
    const EnsArgs E = make_ens(h);
    FusedNodes N{h->nodeset[p][0], h->nodeset[p][1], h->nodeset[p][2], h->nodeset[p][3]};
    if (ens_tiles(h)) {
      tile_args(h, A, p);
      A.rest8_next = h->d_rest8 + (size_t)p * h->sz.nel;  // the first step's own bytes
      BL_LAUNCH(h, "f1_seg", k_f1_seg, (unsigned)h->ntile, kBlock, A, N, E);
    } else {
      BL_LAUNCH(h, "e1_getdt", k_e1_getdt, nblocks(h->sz.nel), kBlock, A, N, E);
    }
  }
  for (int k = 0; k < nsteps; ++k) {
    if (h->timing) {
      if (launch_ensemble_step(h, p)) return BL_FAILURE;
    } else {
      BL_HIP(hipGraphLaunch(h->graph_exec[p], h->stream));
    }
    p ^= 1;
  }
  BL_HIP(hipMemcpyAsync(d, h->d_ts, sizeof(DevTimestep), hipMemcpyDeviceToHost, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  const int taken = d->nstep - start;
  if (d->err == 0 && d->epend != kNoPending) {
    // the smallest instance whose dt selection failed, and how
    d->err = d->epend & 3;
    d->err_idx = d->epend >> 2;
  }
  point_fields_at(h, (h->cur_set + d->flips) & 1);
  // (on the tiling the last step wrote rscratch23/24 itself)
  const bool tiles = ens_tiles(h);
  if (tiles) point_elements_at(h, h->cur_set);
  {
    int nel = h->sz.nel;
    BL_LAUNCH(h, "materialize", k_materialize, nblocks(nel), kBlock, nel, h->I(BL_IELND),
              h->nodeset[0][0], h->nodeset[0][1], h->nodeset[1][0], h->nodeset[1][1],
              (const DevTimestep*)nullptr, h->cur_set, h->R(BL_CNX), h->R(BL_CNY),
              h->R(BL_A1),
              h->R(BL_A2), h->R(BL_A3), h->R(BL_B1), h->R(BL_B2), h->R(BL_B3), h->R(BL_CNWT),
              h->R(BL_ELVOLUME), tiles ? nullptr : (const double*)h->cterm[0],
              tiles ? nullptr : (const double*)h->cterm[1], h->R(BL_RSCRATCH23),
              h->R(BL_RSCRATCH24));
  }
  if (taken > 0 && (dt_hist || idt_hist || ctl_hist)) {
    const size_t n = (size_t)taken * h->ens_n;
    std::vector<double> hist(3 * n);
    BL_HIP(hipMemcpyAsync(hist.data(), h->d_ehist, sizeof(double) * 3 * n,
                          hipMemcpyDeviceToHost, h->stream));
    BL_HIP(hipStreamSynchronize(h->stream));
    for (size_t q = 0; q < n; ++q) {
      if (dt_hist) dt_hist[q] = hist[3 * q];
      if (idt_hist) idt_hist[q] = (int32_t)hist[3 * q + 1];
      if (ctl_hist) ctl_hist[q] = (int32_t)hist[3 * q + 2];
    }
  }
  BL_HIP(hipStreamSynchronize(h->stream));
  if (steps_taken) *steps_taken = taken;
  switch (d->err) {
    case 1:
      fill_err(err, d->err_idx, BL_HALT_SINGLE,
               "ERROR: negative volume in geometry_dr_getgeometry");
      set_error("negative volume in geometry_dr_getgeometry");
      return BL_FAILURE;
    case 2:
      fill_err(err, BL_FAILURE, BL_HALT_SINGLE,
               "ERROR: negative CFL condition in hydro_dr_getdt");
      set_error("negative CFL condition in hydro_dr_getdt (instance " +
                std::to_string(d->err_idx) + ")");
      return BL_FAILURE;
    case 3:
      fill_err(err, BL_FAILURE, BL_HALT_ALL, "ERROR: dt < dt_min");
      set_error("dt < dt_min (instance " + std::to_string(d->err_idx) + ")");
      return BL_FAILURE;
    default:
      return BL_SUCCESS;
  }
}

// The launch-shape autotuner: times the limiter pass (F1L, or E1L for an
// ensemble) of the current state under each candidate shape -- threads per
// workgroup 64/128/256 and workgroups 128..2048 (single problem) or
// elements per workgroup and instance 1024..16384 (ensemble) -- and keeps
// the fastest.  Each trial runs the viscosity pass (F1/E1) that writes the
// at-rest mask (an ensemble's worklists), then the limiter pass between two
// events; neither touches the state a step carries over (the viscosity
// fields they write are saved and restored, the masks and worklist counters
// reset).  The choice changes no result: the limiter's CFL selection is a
// first-index minimum whatever the split.
int bl_autotune(bl_handle_t* h, int32_t reps, bl_tuning_t* out) {
  bl_touch(h);
  if (!h || reps < 1) {
    set_error("bl_autotune: reps >= 1");
    return BL_FAILURE;
  }
  const bool ens = h->ens_n > 0;
  if (!ens && (!h->maps_built || !h->d_eos_itype || !h->d_zdtnotreg || !h->d_outer)) {
    set_error("bl_autotune: mesh maps, regions or EOS not set");
    return BL_FAILURE;
  }
  if (ensure_nodesets(h)) return BL_FAILURE;
  // the device timestep as the last run left it (done set after an error or
  // time_end), put back at the end; the trials run with done and skip clear,
  // as the next run starts
  DevTimestep ts_saved, ts_trial;
  BL_HIP(hipMemcpyAsync(&ts_saved, h->d_ts, sizeof(DevTimestep), hipMemcpyDeviceToHost,
                        h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  ts_trial = ts_saved;
  ts_trial.done = 0;
  ts_trial.skip = 0;
  BL_HIP(hipMemcpyAsync(h->d_ts, &ts_trial, sizeof(DevTimestep), hipMemcpyHostToDevice,
                        h->stream));
  const int nel = h->sz.nel;
  const unsigned nbe = nblocks(nel);
  // save the viscosity outputs the trials overwrite
  double* save = nullptr;
  BL_HIP(hipMalloc(&save, sizeof(double) * 9 * (size_t)std::max(nel, 1)));
  const size_t c4 = sizeof(double) * 4 * (size_t)nel, c1 = sizeof(double) * (size_t)nel;
  hipMemcpyAsync(save, h->R(BL_CNVISCX), c4, hipMemcpyDeviceToDevice, h->stream);
  hipMemcpyAsync(save + 4 * (size_t)nel, h->R(BL_CNVISCY), c4, hipMemcpyDeviceToDevice, h->stream);
  hipMemcpyAsync(save + 8 * (size_t)nel, h->R(BL_ELVISC), c1, hipMemcpyDeviceToDevice, h->stream);
  FusedArgs A = make_args(h);
  A.stop_at_time_end = 0;
  A.time_end = h->cfg.time_end;
  const int p = h->cur_set;
  FusedNodes N{h->nodeset[p][0], h->nodeset[p][1], h->nodeset[p][2], h->nodeset[p][3]};
  EnsArgs E{};
  if (ens) {
    A.einst = h->d_einst;
    A.ninst = h->d_ninst;
    A.inst_dt = h->d_inst_dt;
    A.inst_rdt = h->d_inst_dt + h->ens_n;
  }
  // an ensemble's worklist lengths of the "every element" trials
  std::vector<uint32_t> full(ens ? h->ens_n : 1, (uint32_t)nel);
  if (ens)
    for (int k = 0; k < h->ens_n; ++k) full[k] = h->ens_eoff[k + 1] - h->ens_eoff[k];
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BL_HIP(hipEventCreate(&e0));
  BL_HIP(hipEventCreate(&e1));
  const int blocks[3] = {64, 128, 256};
  const int grids[5] = {128, 256, 512, 1024, 2048};
  const int chunks[5] = {1024, 2048, 4096, 8192, 16384};
  const int def_block = h->tune_lim_block, def_second = ens ? h->tune_ens_chunk : h->tune_lim_grid;
  double tcur[15], tfull[15];
  int cblock[15], csecond[15], trials = 0;
  int rc = BL_SUCCESS;
  for (int b = 0; b < 3 && rc == BL_SUCCESS; ++b)
    for (int g = 0; g < 5 && rc == BL_SUCCESS; ++g) {
      const int block = blocks[b], second = ens ? chunks[g] : grids[g];
      if (ens && ens_set_chunk(h, second)) { rc = BL_FAILURE; break; }
      if (ens) E = make_ens(h);
      // pass 0: the elements in motion in the current state; pass 1: every
      // element in motion (the flow spread everywhere)
      for (int pass = 0; pass < 2 && rc == BL_SUCCESS; ++pass) {
        std::vector<float> t;
        // every element in motion: the single problem's mask all zero, an
        // ensemble's worklist every element (and its counts below)
        if (pass == 1 && !ens)
          hipMemsetAsync(h->d_rest, 0, sizeof(uint64_t) * 4 * (size_t)nbe, h->stream);
        if (pass == 1 && ens)
          hipLaunchKernelGGL(k_work_all, dim3(nbe), dim3(kBlock), 0, h->stream, h->d_work, nel);
        for (int r = 0; r <= reps; ++r) {
          uint32_t* cnt = h->d_cnt;
          if (pass == 0) {
            if (ens) hipMemsetAsync(cnt, 0, sizeof(uint32_t) * full.size(), h->stream);
            if (ens)
              hipLaunchKernelGGL(k_e1_getdt, dim3(nbe), dim3(kBlock), 0, h->stream, A, N, E);
            else
              hipLaunchKernelGGL(k_f1_getdt, dim3(nbe), dim3(kBlock), 0, h->stream, A, N);
          } else if (ens) {
            hipMemcpyAsync(cnt, full.data(), sizeof(uint32_t) * full.size(),
                           hipMemcpyHostToDevice, h->stream);
          }
          hipEventRecord(e0, h->stream);
          if (ens)
            hipLaunchKernelGGL(k_e1_limiter, dim3(h->ens_lblocks), dim3(block), 0, h->stream, A,
                               N, E);
          else
            hipLaunchKernelGGL(k_f1_limiter, dim3(std::min(nbe, (unsigned)second)), dim3(block),
                               0, h->stream, A, N, (int)nbe);
          hipEventRecord(e1, h->stream);
          if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            set_error("bl_autotune: trial launch failed");
            rc = BL_FAILURE;
            break;
          }
          float ms = 0.0f;
          hipEventElapsedTime(&ms, e0, e1);
          if (r > 0) t.push_back(ms);  // the first run warms the caches
        }
        if (rc != BL_SUCCESS) break;
        std::sort(t.begin(), t.end());
        (pass == 0 ? tcur : tfull)[trials] = 1e3 * t[t.size() / 2];
      }
      cblock[trials] = block;
      csecond[trials] = second;
      ++trials;
    }
  // put back what the trials changed
  hipMemcpyAsync(h->R(BL_CNVISCX), save, c4, hipMemcpyDeviceToDevice, h->stream);
  hipMemcpyAsync(h->R(BL_CNVISCY), save + 4 * (size_t)nel, c4, hipMemcpyDeviceToDevice, h->stream);
  hipMemcpyAsync(h->R(BL_ELVISC), save + 8 * (size_t)nel, c1, hipMemcpyDeviceToDevice, h->stream);
  if (ens) hipMemsetAsync(h->d_cnt, 0, sizeof(uint32_t) * h->ens_n, h->stream);
  hipMemcpyAsync(h->d_ts, &ts_saved, sizeof(DevTimestep), hipMemcpyHostToDevice, h->stream);
  clear_rest(h);
  hipStreamSynchronize(h->stream);
  hipFree(save);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != BL_SUCCESS) {
    if (ens) ens_set_chunk(h, def_second);
    return rc;
  }
  // the shape whose worse ratio to the best time of the two passes is
  // smallest (the first such shape on a tie)
  double bc = DBL_MAX, bf = DBL_MAX;
  for (int k = 0; k < trials; ++k) { bc = std::min(bc, tcur[k]); bf = std::min(bf, tfull[k]); }
  int pick = 0, idef = -1;
  double score = DBL_MAX;
  for (int k = 0; k < trials; ++k) {
    const double sk = std::max(tcur[k] / bc, tfull[k] / bf);
    if (sk < score) { score = sk; pick = k; }
    if (cblock[k] == def_block && csecond[k] == def_second) idef = k;
  }
  h->tune_lim_block = cblock[pick];
  if (ens) {
    if (ens_set_chunk(h, csecond[pick])) return BL_FAILURE;
  } else {
    h->tune_lim_grid = csecond[pick];
  }
  invalidate_graph(h);
  if (out) {
    out->limiter_block = cblock[pick];
    out->limiter_grid = ens ? (int32_t)h->ens_lblocks
                            : (int32_t)std::min(nbe, (unsigned)csecond[pick]);
    out->ens_chunk = ens ? csecond[pick] : 0;
    out->trials = trials;
    out->best_us = tcur[pick];
    out->default_us = idef >= 0 ? tcur[idef] : 0.0;
    out->best_full_us = tfull[pick];
    out->default_full_us = idef >= 0 ? tfull[idef] : 0.0;
  }
  return BL_SUCCESS;
}

// counters of the last run call (bl_run_graph or bl_ensemble_run)
int bl_run_stats(bl_handle_t* h, int64_t* moving, int64_t* zeroed) {
  if (!h) return BL_FAILURE;
  std::vector<unsigned long long> c(kStatWords, 0ull);
  if (h->d_vstats) {
    BL_HIP(hipMemcpyAsync(c.data(), h->d_vstats, sizeof(unsigned long long) * kStatWords,
                          hipMemcpyDeviceToHost, h->stream));
    BL_HIP(hipStreamSynchronize(h->stream));
  }
  unsigned long long z = 0;
  for (int s = 0; s < kStatSlots; ++s) z += c[kStatStride * (1 + s)];
  if (moving) *moving = (int64_t)c[0];
  if (zeroed) *zeroed = (int64_t)z;
  return BL_SUCCESS;
}

// a launch shape chosen before (by bl_autotune on an earlier handle of the
// same problem, say) applied without timing
int bl_set_tuning(bl_handle_t* h, const bl_tuning_t* t) {
  bl_touch(h);
  if (!h || !t || t->limiter_block < 64 || t->limiter_block > kBlock ||
      t->limiter_block % 64 != 0) {
    set_error("bl_set_tuning: limiter_block must be 64, 128, 192 or 256");
    return BL_FAILURE;
  }
  if (h->ens_n > 0) {
    if (ens_set_chunk(h, t->ens_chunk)) return BL_FAILURE;
  } else {
    if (t->limiter_grid < 1) {
      set_error("bl_set_tuning: limiter_grid must be >= 1");
      return BL_FAILURE;
    }
    h->tune_lim_grid = t->limiter_grid;
  }
  h->tune_lim_block = t->limiter_block;
  invalidate_graph(h);
  return BL_SUCCESS;
}

// the tile step on (the default) or off
int bl_set_tiling(bl_handle_t* h, int32_t on) {
  bl_touch(h);
  if (!h) {
    set_error("bl_set_tiling: null handle");
    return BL_FAILURE;
  }
  if ((on != 0) != h->tiles_on) {
    h->tiles_on = on != 0;
    invalidate_graph(h);
  }
  return BL_SUCCESS;
}

int bl_get_tiling(bl_handle_t* h, int32_t* on, int32_t* ntile, int32_t* nboundary,
                  int32_t* nentries) {
  if (!h) {
    set_error("bl_get_tiling: null handle");
    return BL_FAILURE;
  }
  const bool built = h->tile_state == 1;
  if (on) *on = (h->tiles_on && built && (fold_f1(h) || h->ens_n > 0)) ? 1 : 0;
  if (ntile) *ntile = built ? h->ntile : 0;
  if (nboundary) *nboundary = 0;  // every element completes in its own tile
  if (nentries) *nentries = built ? h->nentries : 0;
  return BL_SUCCESS;
}

// the launch shape in use
int bl_get_tuning(bl_handle_t* h, bl_tuning_t* t) {
  if (!h || !t) {
    set_error("bl_get_tuning: null argument");
    return BL_FAILURE;
  }
  const bool ens = h->ens_n > 0;
  t->limiter_block = h->tune_lim_block;
  t->limiter_grid = ens ? (int32_t)h->ens_lblocks
                        : (int32_t)std::min(nblocks(h->sz.nel), (unsigned)h->tune_lim_grid);
  t->ens_chunk = ens ? h->tune_ens_chunk : 0;
  t->trials = 0;
  t->best_us = 0.0;
  t->default_us = 0.0;
  t->best_full_us = 0.0;
  t->default_full_us = 0.0;
  return BL_SUCCESS;
}

int bl_run_graph(bl_handle_t* h, int32_t nsteps, int32_t stop_at_time_end, bl_timestep_t* ts,
                 double* dt_hist, int32_t* idt_hist, int32_t* ctl_hist, int32_t* steps_taken,
                 bl_error_t* err) {
  if (!h || !ts) return BL_FAILURE;
  // continue from the last run call's end (see Handle::resumable)
  const bool resume = h->resumable && !h->ptr_out && fold_f1(h);
  h->resumable = false;
  if (h->ens_n > 0) {
    set_error("bl_run_graph: an ensemble handle steps with bl_ensemble_run");
    return BL_FAILURE;
  }
  if (!h->maps_built || !h->d_eos_itype || !h->d_zdtnotreg) {
    set_error("bl_run_graph: mesh maps, regions or EOS not set");
    return BL_FAILURE;
  }
  // split advection, ALE with neither Eulerian nor relaxation mesh motion, or
  // a mesh without the limiter stencil table: host-driven
  if ((h->cfg.zexist &&
       (h->cfg.adv_type != 1 || (!h->cfg.zeul && h->relax_type == BL_RELAX_OFF))) ||
      !h->d_outer)
    return bl_run(h, nsteps, stop_at_time_end, ts, dt_hist, idt_hist, ctl_hist, steps_taken,
                  err);
  fill_err(err, BL_SUCCESS, 0, "");
  if (ensure_nodesets(h) || prepare_tiles(h)) return BL_FAILURE;
  if (h->hist_cap < nsteps || !h->h_pub_hist) {
    // the graphs bake in the history's capacity: size it generously so that
    // runs of different lengths do not re-capture
    const int cap = std::max(std::max(nsteps, (int)h->hist_cap), kMinHistory);
    if (h->d_hist) hipFree(h->d_hist);
    if (h->h_pub_hist) hipHostFree(h->h_pub_hist);
    h->d_hist = nullptr;
    h->h_pub_hist = nullptr;
    BL_HIP(hipMalloc(&h->d_hist, sizeof(double) * 3 * cap));
    BL_HIP(hipHostMalloc(&h->h_pub_hist, sizeof(double) * 3 * cap, hipHostMallocCoherent));
    h->hist_cap = cap;
    invalidate_graph(h);
  }
  // the graphs bake in the node set pointers, which are fixed; each is
  // captured once, before the call's first launch
  if (h->graph_stop_at_end != stop_at_time_end) {
    invalidate_graph(h);
    h->graph_stop_at_end = stop_at_time_end;
  }
  if (h->cfg.zexist && (h->nbfc_stale || !h->d_nbfc) && build_nbfc(h)) return BL_FAILURE;
  // (all of them at once, so that a later call of another length finds
  // its graphs: a capture costs about 0.1 ms)
  const std::vector<int> plan = graph_plan(h->timing ? 0 : nsteps, h->cur_set);
  bool missing = false;
  for (int q : plan) missing = missing || !h->graph_exec[q];
  for (int q = 0; missing && q < kGraphs; ++q)
    if (!h->graph_exec[q] && capture_graph(h, q / (2 * kGraphLogs), (q / 2) % kGraphLogs, q & 1,
                                           stop_at_time_end))
      return BL_FAILURE;
  DevTimestep* d = h->h_ts;
  std::memset(d, 0, sizeof(DevTimestep));
  d->time = ts->time; d->dt = ts->dt; d->dts = ts->dts;
  d->nstep = ts->nstep; d->idtel = ts->idtel; d->sdt = ts->sdt;
  d->last_step = ts->nstep + nsteps;
  d->set0 = h->cur_set;
  if (clear_rest(h, d, resume)) return BL_FAILURE;
  const bool tiles = use_tiles(h);
  int p = h->cur_set;
  // F1 of the first step: on a resumed call the last call's last step left
  // its terms, its at-rest mask and the zeros it stands for
  if (fold_f1(h) && !resume && launch_f1_prologue(h, p, stop_at_time_end)) return BL_FAILURE;
  // with kernel timing on, the same step kernels are launched directly so
  // each launch can be bracketed by HIP events on this stream
  if (h->timing)
    for (int k = 0; k < nsteps; ++k, p ^= 1)
      if (launch_step(h, p, stop_at_time_end)) return BL_FAILURE;
  for (int q : plan) BL_HIP(hipGraphLaunch(h->graph_exec[q], h->stream));
  // everything after the batch is queued behind it with no host round trip:
  // the batch's end (in the last graph, or launched here: the open end of
  // the last step, the geometry at the final node set, which every step
  // whose F3 ran flipped; the device knows how many did), then one read of
  // the timestep, the dt history (at most one entry per step) and the error
  // state
  if (plan.empty() && launch_batch_end(h, h->d_ts, 0)) return BL_FAILURE;
  if (!h->cfg.zexist && mix_cpu_overlay(h, nullptr)) return BL_FAILURE;
  // the batch's end published the timestep and the history (k_publish)
  BL_HIP(hipStreamSynchronize(h->stream));
  *d = *h->h_pub_ts;
  const double* hist = h->h_pub_hist;
  // steps that selected a dt (a failing step included)
  const int taken = d->hist_n;
  for (int k = 0; k < taken; ++k) {
    if (dt_hist) dt_hist[k] = hist[3 * k];
    if (idt_hist) idt_hist[k] = (int32_t)hist[3 * k + 1];
    if (ctl_hist) ctl_hist[k] = (int32_t)hist[3 * k + 2];
  }
  point_fields_at(h, (h->cur_set + d->flips) & 1);
  if (tiles) point_elements_at(h, h->cur_set);
  ts->time = d->time; ts->dt = d->dt; ts->dts = d->dts;
  ts->nstep = d->nstep; ts->idtel = d->idtel; ts->sdt = d->sdt;
  ts->zcorrector = 1;
  if (steps_taken) *steps_taken = taken;
  // every step ran: the next call may continue from here
  h->resumable = d->err == 0 && taken == nsteps && nsteps > 0 && fold_f1(h) && !h->ptr_out;
  switch (d->err) {
    case 1:
      fill_err(err, d->err_idx, BL_HALT_SINGLE,
               "ERROR: negative volume in geometry_dr_getgeometry");
      set_error("negative volume in geometry_dr_getgeometry");
      return BL_FAILURE;
    case 2:
      fill_err(err, BL_FAILURE, BL_HALT_SINGLE,
               "ERROR: negative CFL condition in hydro_dr_getdt");
      set_error("negative CFL condition in hydro_dr_getdt");
      return BL_FAILURE;
    case 3:
      fill_err(err, BL_FAILURE, BL_HALT_ALL, "ERROR: dt < dt_min");
      set_error("dt < dt_min");
      return BL_FAILURE;
    default:
      return BL_SUCCESS;
  }
}

}  // extern "C"
