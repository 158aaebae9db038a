// bl_relax.hip -- node relaxation: the stencil of the mesh generator's SOR
// (setup_mesh_region, setup_mesh.F90:795-877) as Jacobi sweeps on the
// device, used as the ALE step's mesh motion (bl_set_relaxation) and on its
// own (bl_relax_mesh).
//
// Each node's 8-node stencil is a table built once on the host from the
// connectivity: (8, nnd) int32, 0-based, in the order of the weights r1..r8
// of setup_mesh.F90:848-857 -- (lp,km) (lp,kk) (lp,kp) (ll,km) (ll,kp)
// (lm,km) (lm,kk) (lm,kp) -- and -1 for nodes that are not relaxed
// (boundary nodes, or fewer or more than four elements around them).  A
// sweep is one thread per node: 32 B of table, the node and its eight
// neighbours' coordinates (shared with the neighbouring threads through the
// cache), 16 B out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "bl_internal.h"
#include "bl_math.h"
#include "bl_relax.h"

namespace bl {

#define RELAX_GATE(ts) \
  if (((ts) && ((ts)->skip || !(ts)->ale_on)) || (R.only_if && *R.only_if == 0)) return

__global__ void k_relax_sweep(RelaxArgs R, const double* __restrict__ xi,
                              const double* __restrict__ yi, double* __restrict__ xo,
                              double* __restrict__ yo, const DevTimestep* ts) {
  RELAX_GATE(ts);
  const int stride = gridDim.x * blockDim.x;
  for (int n = xcd_index(); n < R.nnd; n += stride) {
    double a, b;
    relax_node(R, n, xi, yi, a, b);
    xo[n] = a;
    yo[n] = b;
  }
}

// the last sweep of the ALE mesh motion: w = (x_R - x)/dt, x the Lagrangian
// positions (xl, yl), x_R this sweep's result from (xi, yi)
__global__ void k_relax_velocity(RelaxArgs R, const double* __restrict__ xl,
                                 const double* __restrict__ yl, const double* __restrict__ xi,
                                 const double* __restrict__ yi, double* __restrict__ wu,
                                 double* __restrict__ wv, double dt, const DevTimestep* ts) {
  RELAX_GATE(ts);
  if (ts) dt = ts->dt;
  const int stride = gridDim.x * blockDim.x;
  for (int n = xcd_index(); n < R.nnd; n += stride) {
    double a, b;
    relax_node(R, n, xi, yi, a, b);
    wu[n] = (a - xl[n]) / dt;
    wv[n] = (b - yl[n]) / dt;
  }
}

}  // namespace bl

using namespace bl;

namespace {

// the stencil table of every node (same frame rule as the C-ABI states):
// per node, its elements as (element, corner) in ascending element order;
// e0 the first with the node at corner 1, else the first; then walk
// counter-clockwise, the next element being the one whose corner after the
// node is the current element's corner before it
int build_ring(bl_handle_t* h) {
  const int nel = h->sz.nel, nnd = h->sz.nnd;
  std::vector<int32_t> ielnd(4 * (size_t)nel), indtype(nnd);
  if (bl_download(h, BL_IELND, ielnd.data(), 4 * (int64_t)nel) ||
      bl_download(h, BL_INDTYPE, indtype.data(), nnd))
    return BL_FAILURE;
  std::vector<int32_t> first(nnd + 1, 0), slots(4 * (size_t)nel);
  for (int k = 0; k < 4 * nel; ++k) {
    const int n = ielnd[k] - 1;
    if (n < 0 || n >= nnd) { set_error("relaxation: ielnd out of range"); return BL_FAILURE; }
    first[n + 1] += 1;
  }
  for (int n = 0; n < nnd; ++n) first[n + 1] += first[n];
  std::vector<int32_t> pos(first.begin(), first.end() - 1);
  for (int k = 0; k < 4 * nel; ++k) slots[pos[ielnd[k] - 1]++] = k;  // k = 4*e + corner
  auto at = [&](int slot, int d) { return ielnd[4 * (size_t)(slot >> 2) + ((slot + d) & 3)] - 1; };
  std::vector<int32_t> ring(8 * (size_t)std::max(nnd, 1), -1);
  for (int n = 0; n < nnd; ++n) {
    const int b = first[n], e = first[n + 1];
    if (indtype[n] <= 0 || e - b != 4) continue;
    int cur = slots[b];
    for (int k = b; k < e; ++k)
      if ((slots[k] & 3) == 0) { cur = slots[k]; break; }
    // counter-clockwise from +k: (ll,kp) (lp,kp) (lp,kk) (lp,km) (ll,km)
    // (lm,km) (lm,kk) (lm,kp)
    int ccw[8];
    bool seen[4] = {false, false, false, false};
    bool ok = true;
    for (int q = 0; q < 4 && ok; ++q) {
      if (q > 0) {
        const int want = at(cur, 3);
        int nxt = -1;
        for (int k = b; k < e; ++k)
          if (!seen[k - b] && at(slots[k], 1) == want) { nxt = k; break; }
        if (nxt < 0) { ok = false; break; }
        cur = slots[nxt];
        seen[nxt - b] = true;
      } else {
        for (int k = b; k < e; ++k)
          if (slots[k] == cur) seen[k - b] = true;
      }
      if (q == 0) ccw[0] = at(cur, 1);
      ccw[2 * q + 1] = at(cur, 2);
      if (q < 3) ccw[2 * q + 2] = at(cur, 3);
    }
    if (!ok || at(cur, 3) != ccw[0]) continue;
    int32_t* r = &ring[8 * (size_t)n];
    r[0] = ccw[3]; r[1] = ccw[2]; r[2] = ccw[1]; r[3] = ccw[4];
    r[4] = ccw[0]; r[5] = ccw[5]; r[6] = ccw[6]; r[7] = ccw[7];
  }
  if (h->d_ring) hipFree(h->d_ring);
  h->d_ring = nullptr;
  BL_HIP(hipMalloc(&h->d_ring, sizeof(int32_t) * ring.size()));
  BL_HIP(hipMemcpy(h->d_ring, ring.data(), sizeof(int32_t) * ring.size(), hipMemcpyHostToDevice));
  // the relaxed nodes and the others, each ascending (the learned
  // relaxation's work lists)
  std::vector<int32_t> rlist, zlist;
  for (int n = 0; n < nnd; ++n) (ring[8 * (size_t)n] >= 0 ? rlist : zlist).push_back(n);
  for (int k = 0; k < 2; ++k) {
    int32_t*& d = k == 0 ? h->d_rlist : h->d_zlist;
    const std::vector<int32_t>& v = k == 0 ? rlist : zlist;
    if (d) hipFree(d);
    d = nullptr;
    BL_HIP(hipMalloc(&d, sizeof(int32_t) * std::max<size_t>(v.size(), 1)));
    if (!v.empty())
      BL_HIP(hipMemcpy(d, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice));
  }
  h->nrelaxed = (int32_t)rlist.size();
  h->nzero = (int32_t)zlist.size();
  return BL_SUCCESS;
}

int ensure_tmp(bl_handle_t* h) {
  if (h->relax_tmp[0]) return BL_SUCCESS;
  for (int k = 0; k < 4; ++k)
    BL_HIP(hipMalloc(&h->relax_tmp[k], sizeof(double) * std::max(h->sz.nnd, 1)));
  return BL_SUCCESS;
}

bool args_ok(int32_t type, int32_t nsweep, double om, double tol) {
  return type >= BL_RELAX_OFF && type <= BL_RELAX_EQUI && nsweep >= 1 && om > 0.0 && tol >= 0.0;
}

RelaxArgs make_relax(bl_handle_t* h, int type, double om, double tol) {
  RelaxArgs R;
  R.nnd = h->sz.nnd;
  R.type = type;
  R.om = om;
  R.tol = tol;
  R.zerocut = h->cfg.zerocut;
  R.ring = h->d_ring;
  R.only_if = nullptr;
  return R;
}

}  // namespace

namespace bl {

RelaxArgs relax_args(Handle* h) {
  return make_relax(static_cast<bl_handle_t*>(h), h->relax_type, h->relax_om, h->relax_tol);
}

int relax_rebuild(Handle* h) {
  if (h->relax_type == BL_RELAX_OFF && !h->d_ring) return BL_SUCCESS;
  return build_ring(static_cast<bl_handle_t*>(h));
}

// the ALE mesh velocity of the relaxation into (wu, wv) from the positions
// (x, y); with ts the kernels are gated like the other ALE kernels and dt
// is the device timestep's
int relax_velocity(Handle* hh, const double* x, const double* y, double* wu, double* wv,
                   double dt, const DevTimestep* ts, const int32_t* only_if) {
  bl_handle_t* h = static_cast<bl_handle_t*>(hh);
  const int nnd = h->sz.nnd;
  if (!h->d_ring || !h->relax_tmp[0]) {
    set_error("relaxation: stencil table not built (bl_set_relaxation)");
    return BL_FAILURE;
  }
  RelaxArgs R = make_relax(h, h->relax_type, h->relax_om, h->relax_tol);
  R.only_if = only_if;
  // a gated fallback mostly does nothing: a small grid (each thread a few
  // nodes when it runs) keeps the empty launches cheap
  const unsigned grid = only_if ? std::min(nblocks(nnd), kGatedBlocks) : nblocks(nnd);
  const double* xi = x;
  const double* yi = y;
  for (int k = 0; k + 1 < h->relax_nsweep; ++k) {
    double* xo = h->relax_tmp[2 * (k & 1)];
    double* yo = h->relax_tmp[2 * (k & 1) + 1];
    BL_LAUNCH(h, "relax_sweep", k_relax_sweep, grid, kBlock, R, xi, yi, xo, yo, ts);
    xi = xo;
    yi = yo;
  }
  BL_LAUNCH(h, "relax_velocity", k_relax_velocity, grid, kBlock, R, x, y, xi, yi, wu, wv, dt, ts);
  return BL_SUCCESS;
}

}  // namespace bl

extern "C" {

int bl_set_relaxation(bl_handle_t* h, int32_t type, int32_t nsweep, double om, double tol) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  if (!args_ok(type, nsweep, om, tol)) {
    set_error("bl_set_relaxation: type 0..3, nsweep >= 1, om > 0 and tol >= 0");
    return BL_FAILURE;
  }
  BL_HIP(hipStreamSynchronize(h->stream));
  invalidate_graph(h);
  h->relax_type = type;
  h->relax_nsweep = nsweep;
  h->relax_om = om;
  h->relax_tol = tol;
  if (type == BL_RELAX_OFF) return BL_SUCCESS;
  if (build_ring(h) || ensure_tmp(h)) return BL_FAILURE;
  return BL_SUCCESS;
}

int bl_relax_mesh(bl_handle_t* h, int32_t type, int32_t nsweep, double om, double tol) {
  bl_touch(h);
  if (!h) return BL_FAILURE;
  if (!args_ok(type, nsweep, om, tol) || type == BL_RELAX_OFF) {
    set_error("bl_relax_mesh: type 1..3, nsweep >= 1, om > 0 and tol >= 0");
    return BL_FAILURE;
  }
  if (build_ring(h) || ensure_tmp(h)) return BL_FAILURE;
  const int nnd = h->sz.nnd;
  RelaxArgs R = make_relax(h, type, om, tol);
  double* x = h->R(BL_NDX);
  double* y = h->R(BL_NDY);
  const double* xi = x;
  const double* yi = y;
  for (int k = 0; k < nsweep; ++k) {
    const bool last = k + 1 == nsweep;
    double* xo = last ? h->relax_tmp[2] : h->relax_tmp[2 * (k & 1)];
    double* yo = last ? h->relax_tmp[3] : h->relax_tmp[2 * (k & 1) + 1];
    if (last && xi == h->relax_tmp[2]) { xo = h->relax_tmp[0]; yo = h->relax_tmp[1]; }
    BL_LAUNCH(h, "relax_sweep", k_relax_sweep, nblocks(nnd), kBlock, R, xi, yi, xo, yo,
              (const DevTimestep*)nullptr);
    xi = xo;
    yi = yo;
  }
  BL_HIP(hipMemcpyAsync(x, xi, sizeof(double) * nnd, hipMemcpyDeviceToDevice, h->stream));
  BL_HIP(hipMemcpyAsync(y, yi, sizeof(double) * nnd, hipMemcpyDeviceToDevice, h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  return BL_SUCCESS;
}

}  // extern "C"
