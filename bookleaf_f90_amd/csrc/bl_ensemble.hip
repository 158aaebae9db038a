// bl_ensemble.hip -- ensembles of independent problems on one device: mesh
// generation and initial conditions of every instance on the device, and
// the handle the ensemble step (bl_ensemble_run, bl_fused.hip) runs on.
//
// The ensemble's mesh is the disjoint union of the instances' structured
// meshes.  Instance k owns elements [eoff[k], eoff[k+1]) and nodes
// [noff[k], noff[k+1]); inside an instance the numbering is that of a
// structured nx x ny block (element i + j*nx with corners (i,j), (i+1,j),
// (i+1,j+1), (i,j+1); node i + j*(nx+1)), the layout the mesh generator's
// setup_mesh_transfer gives a single LIN1 region.  Every table the step
// reads -- ielnd, ielel/ielfc, node types, the three node-to-corner CSR
// orders, the limiter's outer-node table -- follows from (i, j) in closed
// form, so the whole set-up is a few grid kernels with no host pass over the
// mesh.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bl_fused.h"
#include "bl_internal.h"
#include "bl_math.h"

namespace bl {

struct GenArgs {
  int ninst, nel, nnd;
  const int32_t* eoff;
  const int32_t* noff;
  const bl_instance_t* inst;
  int32_t *ielnd, *ielel, *ielfc, *ielreg, *ielmat, *ielsort1, *ielsort2, *indtype;
  double *rho, *en, *ndx, *ndy, *ndu, *ndv;
  uint16_t *einst, *ninst_of;
  int32_t *hyd_ptr, *hyd_idx, *ale_ptr, *ale_idx, *el_ptr, *el_idx, *outer;
};

// the instance k with off[k] <= idx < off[k+1]
__device__ __forceinline__ int find_inst(const int32_t* __restrict__ off, int n, int idx) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// node coordinates of structured_mesh on [0, L] x [0, H]: (L*i)/nx
__device__ __forceinline__ double coord(double len, int i, int n) {
  return (len * (double)i) / (double)n;
}

__global__ void k_gen_elements(GenArgs G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G.nel) return;
  const int k = find_inst(G.eoff, G.ninst, e);
  const bl_instance_t& P = G.inst[k];
  const int nx = P.nx, ny = P.ny;
  const int le = e - G.eoff[k];
  const int i = le % nx, j = le / nx;
  const int n00 = G.noff[k] + i + j * (nx + 1);
  const int nd[4] = {n00, n00 + 1, n00 + 1 + (nx + 1), n00 + (nx + 1)};
  reinterpret_cast<int4*>(G.ielnd)[e] = make_int4(nd[0] + 1, nd[1] + 1, nd[2] + 1, nd[3] + 1);
  // face f joins corners f and f+1; a boundary face names the element
  // itself (utils_kn_corrconn); the neighbour's face back is f+2
  const int el1 = e + 1;
  reinterpret_cast<int4*>(G.ielel)[e] =
      make_int4(j > 0 ? el1 - nx : el1, i < nx - 1 ? el1 + 1 : el1,
                j < ny - 1 ? el1 + nx : el1, i > 0 ? el1 - 1 : el1);
  reinterpret_cast<int4*>(G.ielfc)[e] = make_int4(3, 4, 1, 2);
  G.ielsort1[e] = el1;
  G.ielsort2[e] = el1;
  G.ielmat[e] = k + 1;
  G.einst[e] = (uint16_t)k;
  const double x[4] = {coord(P.length, i, nx), coord(P.length, i + 1, nx),
                       coord(P.length, i + 1, nx), coord(P.length, i, nx)};
  const double y[4] = {coord(P.height, j, ny), coord(P.height, j, ny),
                       coord(P.height, j + 1, ny), coord(P.height, j + 1, ny)};
  if (P.kind == 0) {
    // Sod: the centroid's side of length/2 (the mean of the corners)
    const double xc = (((x[0] + x[1]) + x[2]) + x[3]) / 4.0;
    const bool left = xc < 0.5 * P.length;
    G.ielreg[e] = 2 * k + (left ? 1 : 2);
    G.rho[e] = left ? P.rho_l : P.rho_r;
    G.en[e] = left ? P.e_l : P.e_r;
  } else {
    // Sedov: e_l deposited in the element at the origin, e = E/(rho V1)
    // with V1 by the operation order of geometry_kn_getiso/getvolume
    G.ielreg[e] = 2 * k + 1;
    G.rho[e] = P.rho_l;
    double en = 0.0;
    if (le == 0) {
      const double a1 = 0.25 * (-x[0] + x[1] + x[2] - x[3]);
      const double a3 = 0.25 * (-x[0] - x[1] + x[2] + x[3]);
      const double b1 = 0.25 * (-y[0] + y[1] + y[2] - y[3]);
      const double b3 = 0.25 * (-y[0] - y[1] + y[2] + y[3]);
      const double v1 = 4.0 * (a1 * b3 - a3 * b1);
      en = P.e_l / (P.rho_l * v1);
    }
    G.en[e] = en;
  }
}

// nodes: coordinates, zero velocity, node types of reflecting walls
// (SLIPY bottom and top -> -2, SLIPX left and right -> -1, corners -3), and
// the three node-to-corner CSR orders: hydro (corner j, element), ALE and
// element order (element, corner).  Node (i,j) is corner 1 of (i,j), 2 of
// (i-1,j), 3 of (i-1,j-1) and 4 of (i,j-1), those that exist.
__global__ void k_gen_nodes(GenArgs G) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.nnd) return;
  const int k = find_inst(G.noff, G.ninst, n);
  const bl_instance_t& P = G.inst[k];
  const int nx = P.nx, ny = P.ny;
  const int ln = n - G.noff[k];
  const int i = ln % (nx + 1), j = ln / (nx + 1);
  G.ndx[n] = coord(P.length, i, nx);
  G.ndy[n] = coord(P.height, j, ny);
  G.ndu[n] = 0.0;
  G.ndv[n] = 0.0;
  G.ninst_of[n] = (uint16_t)k;
  const bool xb = (i == 0 || i == nx), yb = (j == 0 || j == ny);
  G.indtype[n] = (xb && yb) ? -3 : (yb ? -2 : (xb ? -1 : 1));
  // CSR: corners of the nodes before this one
  const int aj = (j > 0) + (j < ny);
  const int p = 4 * G.eoff[k] + (j >= 1 ? (2 * j - 1) * 2 * nx : 0) +
                aj * (i >= 1 ? 2 * i - 1 : 0);
  G.hyd_ptr[n] = p;
  G.ale_ptr[n] = p;
  G.el_ptr[n] = p;
  if (n == G.nnd - 1) {
    G.hyd_ptr[G.nnd] = 4 * G.nel;
    G.ale_ptr[G.nnd] = 4 * G.nel;
    G.el_ptr[G.nnd] = 4 * G.nel;
  }
  const int eb = G.eoff[k];
  const bool h00 = i < nx && j < ny, h10 = i > 0 && j < ny, h11 = i > 0 && j > 0,
             h01 = i < nx && j > 0;
  const int c00 = 4 * (eb + i + j * nx), c10 = 4 * (eb + i - 1 + j * nx) + 1,
            c11 = 4 * (eb + i - 1 + (j - 1) * nx) + 2, c01 = 4 * (eb + i + (j - 1) * nx) + 3;
  int q = p;
  if (h00) G.hyd_idx[q++] = c00;
  if (h10) G.hyd_idx[q++] = c10;
  if (h11) G.hyd_idx[q++] = c11;
  if (h01) G.hyd_idx[q++] = c01;
  q = p;
  if (h11) { G.ale_idx[q] = c11; G.el_idx[q] = c11; ++q; }
  if (h01) { G.ale_idx[q] = c01; G.el_idx[q] = c01; ++q; }
  if (h10) { G.ale_idx[q] = c10; G.el_idx[q] = c10; ++q; }
  if (h00) { G.ale_idx[q] = c00; G.el_idx[q] = c00; ++q; }
}

// the limiter's outer-node table from ielnd, ielel, ielfc: for direction
// is of element e, the far nodes of the neighbours across faces is and
// is+2, in the order F1 reads them (a boundary side takes e's own nodes)
__global__ void k_gen_outer(GenArgs G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G.nel) return;
  auto node = [&](int el, int c) { return G.ielnd[4 * el + (c & 3)]; };
  for (int is = 0; is < 2; ++is) {
    int32_t* o = G.outer + 8 * e + 4 * is;
    const int in1 = G.ielel[4 * e + is] - 1, in2 = G.ielel[4 * e + is + 2] - 1;
    const int f1 = G.ielfc[4 * e + is] - 1, f2 = G.ielfc[4 * e + is + 2] - 1;
    if (in1 == e) { o[0] = node(e, is); o[1] = node(e, is + 1); }
    else { o[0] = node(in1, f1 + 2); o[1] = node(in1, f1 + 3); }
    if (in2 == e) { o[2] = node(e, is + 3); o[3] = node(e, is + 2); }
    else { o[2] = node(in2, f2 + 3); o[3] = node(in2, f2 + 2); }
  }
}

}  // namespace bl

using namespace bl;

namespace {

template <typename T>
int dev_alloc(T** p, size_t n) {
  if (hipMalloc(p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
    set_error("bl_ensemble_create: out of device memory");
    return BL_FAILURE;
  }
  return BL_SUCCESS;
}

}  // namespace

namespace bl {

// the limiter pass's workgroups per instance, ceil(nel_k / chunk), their
// offsets lboff on the device, and room for their partials
int ens_set_chunk(Handle* hh, int chunk) {
  bl_handle_t* h = static_cast<bl_handle_t*>(hh);
  const int ninst = h->ens_n;
  if (chunk < 64) {
    set_error("ensemble limiter chunk must be >= 64 elements");
    return BL_FAILURE;
  }
  std::vector<int32_t> lboff(ninst + 1, 0);
  for (int k = 0; k < ninst; ++k)
    lboff[k + 1] = lboff[k] + (h->ens_eoff[k + 1] - h->ens_eoff[k] + chunk - 1) / chunk;
  const size_t nb = nblocks(h->sz.nel);
  if (lboff[ninst] > h->ens_lblocks_cap) {
    hipFree(h->d_epart);
    hipFree(h->d_eparti);
    h->d_epart = nullptr;
    h->d_eparti = nullptr;
    h->ens_lblocks_cap = 0;
    BL_HIP(hipMalloc(&h->d_epart, sizeof(double) * (4 * nb + (size_t)lboff[ninst])));
    BL_HIP(hipMalloc(&h->d_eparti, sizeof(int32_t) * (4 * nb + (size_t)lboff[ninst])));
    h->ens_lblocks_cap = lboff[ninst];
  }
  BL_HIP(hipMemcpy(h->d_lboff, lboff.data(), sizeof(int32_t) * (ninst + 1),
                   hipMemcpyHostToDevice));
  h->ens_lblocks = lboff[ninst];
  h->tune_ens_chunk = chunk;
  invalidate_graph(h);
  return BL_SUCCESS;
}

}  // namespace bl

extern "C" {

int bl_ensemble_create(const bl_instance_t* inst, int32_t ninst, const bl_config_t* cfg,
                       bl_handle_t** out) {
  if (!inst || !cfg || !out || ninst < 1 || ninst > 65535) {
    set_error("bl_ensemble_create: 1..65535 instances and a configuration are needed");
    return BL_FAILURE;
  }
  if (cfg->zexist) {
    set_error("bl_ensemble_create: ensembles are Lagrangian (zexist = 0)");
    return BL_FAILURE;
  }
  std::vector<int32_t> eoff(ninst + 1, 0), noff(ninst + 1, 0);
  int64_t ne = 0, nn = 0;
  bool zhg = false;
  for (int k = 0; k < ninst; ++k) {
    const bl_instance_t& P = inst[k];
    if (P.nx < 2 || P.ny < 2 || (int64_t)P.nx * P.ny < kBlock || (P.kind != 0 && P.kind != 1) ||
        !(P.length > 0.0) || !(P.height > 0.0)) {
      set_error("bl_ensemble_create: instance " + std::to_string(k) +
                ": nx, ny >= 2, nx*ny >= 256, kind 0 or 1, positive extent");
      return BL_FAILURE;
    }
    eoff[k] = (int32_t)ne;
    noff[k] = (int32_t)nn;
    ne += (int64_t)P.nx * P.ny;
    nn += (int64_t)(P.nx + 1) * (P.ny + 1);
    if (4 * ne >= INT32_MAX || nn >= INT32_MAX) {
      set_error("bl_ensemble_create: more than 2^31 corners");
      return BL_FAILURE;
    }
    zhg = zhg || P.kappa > 0.0;
  }
  eoff[ninst] = (int32_t)ne;
  noff[ninst] = (int32_t)nn;
  bl_sizes_t sz{};
  sz.nel = sz.nel1 = sz.nel2 = (int32_t)ne;
  sz.nnd = sz.nnd1 = sz.nnd2 = (int32_t)nn;
  sz.nsz = (int32_t)std::max(ne, nn);
  sz.nmat = ninst;
  sz.nreg = 2 * ninst;
  sz.ncp = 0; sz.nmx = 0; sz.mcp = 1; sz.mmx = 1;
  bl_handle_t* h = nullptr;
  if (bl_create(&sz, &h) != BL_SUCCESS) return BL_FAILURE;
  auto fail = [&]() { bl_destroy(h); return BL_FAILURE; };
  bl_config_t c = *cfg;
  c.zhg = zhg ? 1 : 0;
  c.zsp = 0;
  if (bl_set_config(h, &c)) return fail();
  std::vector<double> kap(2 * ninst), pme(2 * ninst, 0.0);
  std::vector<int32_t> zdt(2 * ninst, 0), zml(2 * ninst, 0);
  for (int k = 0; k < ninst; ++k) kap[2 * k] = kap[2 * k + 1] = inst[k].kappa;
  if (bl_set_regions(h, 2 * ninst, kap.data(), pme.data(), zdt.data(), zml.data()))
    return fail();
  std::vector<int32_t> itype(ninst + 1, 0);
  std::vector<double> param((size_t)BL_NEOSPARAM * (ninst + 1), 0.0);
  for (int k = 0; k < ninst; ++k) {
    itype[k + 1] = 1;
    param[(size_t)BL_NEOSPARAM * (k + 1)] = inst[k].gamma;
  }
  if (bl_set_eos(h, ninst, itype.data(), param.data())) return fail();

  // ensemble tables
  h->ens_n = ninst;
  h->ens_eoff = eoff;
  h->ens_noff = noff;
  int32_t* d_noff = nullptr;
  bl_instance_t* d_inst = nullptr;
  const size_t nel = ne, nnd = nn, nb = nblocks((int)ne);
  if (dev_alloc(&h->d_eoff, ninst + 1) || dev_alloc(&d_noff, ninst + 1) ||
      dev_alloc(&d_inst, ninst) || dev_alloc(&h->d_einst, nel) || dev_alloc(&h->d_ninst, nnd) ||
      dev_alloc(&h->d_hyd_ptr, nnd + 1) || dev_alloc(&h->d_hyd_idx, 4 * nel) ||
      dev_alloc(&h->d_ale_ptr, nnd + 1) || dev_alloc(&h->d_ale_idx, 4 * nel) ||
      dev_alloc(&h->d_el_ptr, nnd + 1) || dev_alloc(&h->d_el_idx, 4 * nel) ||
      dev_alloc(&h->d_outer, 8 * nel) || dev_alloc((InstTs**)&h->d_inst, ninst) ||
      dev_alloc(&h->d_inst_dt, 2 * (size_t)ninst) || dev_alloc(&h->d_cnt, ninst) ||
      dev_alloc(&h->d_lboff, ninst + 1) || ens_set_chunk(h, h->tune_ens_chunk)) {
    hipFree(d_noff); hipFree(d_inst);
    return fail();
  }
  hipMemcpy(h->d_eoff, eoff.data(), sizeof(int32_t) * (ninst + 1), hipMemcpyHostToDevice);
  hipMemcpy(d_noff, noff.data(), sizeof(int32_t) * (ninst + 1), hipMemcpyHostToDevice);
  hipMemcpy(d_inst, inst, sizeof(bl_instance_t) * ninst, hipMemcpyHostToDevice);
  std::vector<InstTs> its(ninst);
  for (auto& t : its) {
    t.time = c.time_start; t.dt = c.dt_initial; t.nstep = 0; t.idtel = 0;
    t.sdt = BL_DT_INITIAL; t.pad = 0;
  }
  // each instance's dt, then its 1/dt (forceHG's factor, formed with the dt)
  std::vector<double> dts(2 * (size_t)ninst, c.dt_initial);
  for (int k = 0; k < ninst; ++k) dts[ninst + k] = 1.0 / c.dt_initial;
  hipMemcpy(h->d_inst, its.data(), sizeof(InstTs) * ninst, hipMemcpyHostToDevice);
  hipMemcpy(h->d_inst_dt, dts.data(), sizeof(double) * 2 * ninst, hipMemcpyHostToDevice);
  hipMemset(h->d_cnt, 0, sizeof(uint32_t) * ninst);

  GenArgs G{};
  G.ninst = ninst; G.nel = (int)ne; G.nnd = (int)nn;
  G.eoff = h->d_eoff; G.noff = d_noff; G.inst = d_inst;
  G.ielnd = h->I(BL_IELND); G.ielel = h->I(BL_IELEL); G.ielfc = h->I(BL_IELFC);
  G.ielreg = h->I(BL_IELREG); G.ielmat = h->I(BL_IELMAT);
  G.ielsort1 = h->I(BL_IELSORT1); G.ielsort2 = h->I(BL_IELSORT2); G.indtype = h->I(BL_INDTYPE);
  G.rho = h->R(BL_ELDENSITY); G.en = h->R(BL_ELENERGY);
  G.ndx = h->R(BL_NDX); G.ndy = h->R(BL_NDY); G.ndu = h->R(BL_NDU); G.ndv = h->R(BL_NDV);
  G.einst = h->d_einst; G.ninst_of = h->d_ninst;
  G.hyd_ptr = h->d_hyd_ptr; G.hyd_idx = h->d_hyd_idx; G.ale_ptr = h->d_ale_ptr;
  G.ale_idx = h->d_ale_idx; G.el_ptr = h->d_el_ptr; G.el_idx = h->d_el_idx;
  G.outer = h->d_outer;
  hipLaunchKernelGGL(k_gen_elements, dim3(nblocks((int)ne)), dim3(kBlock), 0, h->stream, G);
  hipLaunchKernelGGL(k_gen_nodes, dim3(nblocks((int)nn)), dim3(kBlock), 0, h->stream, G);
  hipLaunchKernelGGL(k_gen_outer, dim3(nblocks((int)ne)), dim3(kBlock), 0, h->stream, G);
  hipError_t le = hipGetLastError();
  hipError_t se = hipStreamSynchronize(h->stream);
  hipFree(d_noff);
  hipFree(d_inst);
  if (le != hipSuccess || se != hipSuccess) {
    set_error("bl_ensemble_create: generation failed");
    return fail();
  }
  h->maps_built = true;
  // the device timestep of the whole ensemble
  DevTimestep* d = h->h_ts;
  std::memset(d, 0, sizeof(DevTimestep));
  d->time = c.time_start; d->dt = c.dt_initial; d->dts = c.dt_initial;
  hipMemcpy(h->d_ts, d, sizeof(DevTimestep), hipMemcpyHostToDevice);
  // inf_init's state set-up on the generated mesh
  bl_error_t err;
  if (bl_geometry_getgeometry(h, &err) || bl_init_state(h) || bl_hydro_init(h)) return fail();
  *out = h;
  return BL_SUCCESS;
}

int bl_ensemble_offsets(bl_handle_t* h, int32_t* eoff, int32_t* noff) {
  if (!h || h->ens_n <= 0) { set_error("bl_ensemble_offsets: not an ensemble"); return BL_FAILURE; }
  if (eoff) std::copy(h->ens_eoff.begin(), h->ens_eoff.end(), eoff);
  if (noff) std::copy(h->ens_noff.begin(), h->ens_noff.end(), noff);
  return BL_SUCCESS;
}

int bl_ensemble_times(bl_handle_t* h, double* time, double* dt, int32_t* nstep) {
  if (!h || h->ens_n <= 0) { set_error("bl_ensemble_times: not an ensemble"); return BL_FAILURE; }
  std::vector<InstTs> its(h->ens_n);
  BL_HIP(hipMemcpyAsync(its.data(), h->d_inst, sizeof(InstTs) * h->ens_n, hipMemcpyDeviceToHost,
                        h->stream));
  BL_HIP(hipStreamSynchronize(h->stream));
  for (int k = 0; k < h->ens_n; ++k) {
    if (time) time[k] = its[k].time;
    if (dt) dt[k] = its[k].dt;
    if (nstep) nstep[k] = its[k].nstep;
  }
  return BL_SUCCESS;
}

}  // extern "C"
