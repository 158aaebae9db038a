// bl_tiles.hip -- the element tiling of the tile step (host side).
//
// The tile step fuses the fused step's F2, F3 and F4 for a pure-cell
// Lagrangian problem: one workgroup owns a tile of elements.  It computes
// the predictor and the corrector corner terms (F2's work, kept in LDS) of
// its own elements and of the ring of elements around their nodes, sums
// them for every node of its own elements in the reference's scatter order
// (F3's work), and then the corrector state of its own elements (F4's
// work), whose nodes it now holds.  Every element is completed by the one
// tile that owns it, so the F2->F3->F4 corner terms never reach HBM and no
// second pass is needed.  A node that belongs to several tiles is computed
// by each of them from the same inputs (the same bits) and written by one:
// the tile owning the node's first incident element.
//
// The reference sums a node's corner terms in (corner jj, element in
// ielsort1 order) order (hydro_kn_getacceleration.f90:53-72).  When no node
// is corner jj of two elements, that order is simply jj = 1..4 over the
// node's incident corners, which a tile reproduces from one LDS slot per
// (node, jj).  A mesh where some node is corner jj of two elements keeps the
// three-kernel step.
//
// Tiles come from recursive coordinate bisection of the element centroids
// when the tiling is built (setup_rcb.f90 cuts the reference's partitions
// the same way; see Rcb below).  The positions only shape the tiles: any
// tiling gives the same results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "bl_internal.h"
#include "bl_tiles.h"

namespace bl {

namespace {

// Recursive coordinate bisection of the elements by their centroids: each
// cut along the longer side of the element set's bounding box, at the share
// that keeps the tile counts of the two sides proportional, down to about
// kTileTarget elements; a leaf whose nodes or the elements around them do
// not fit one workgroup is cut again.
struct Rcb {
  const std::vector<double>& cx;
  const std::vector<double>& cy;
  const std::vector<int32_t>& ielnd;     // 4 per element, 1-based
  const std::vector<int32_t>& adj_ptr;  // node -> incident elements
  const std::vector<int32_t>& adj;
  std::vector<int32_t> idx;              // element ids, cut in place
  std::vector<int32_t> nstamp;           // per node: last count that saw it
  std::vector<int32_t> estamp;           // per element: the same
  std::vector<std::pair<int, int>> leaves;
  int counter = 0;

  // the nodes of idx[lo, hi) fit the tile's node slots and the elements
  // around them one workgroup
  bool fits(int lo, int hi) {
    ++counter;
    int nn = 0, ne = 0;
    for (int k = lo; k < hi; ++k)
      for (int j = 0; j < 4; ++j) {
        const int v = ielnd[4 * (size_t)idx[k] + j] - 1;
        if (nstamp[v] == counter) continue;
        nstamp[v] = counter;
        if (++nn > kTileNodes) return false;
        for (int q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q)
          if (estamp[adj[q]] != counter) {
            estamp[adj[q]] = counter;
            if (++ne > kTileEls) return false;
          }
      }
    return true;
  }

  void cut(int lo, int hi, int k) {
    if (k <= 1) {
      if (hi - lo < 2 || fits(lo, hi)) {
        leaves.emplace_back(lo, hi);
        return;
      }
      k = 2;
    }
    double x0 = cx[idx[lo]], x1 = x0, y0 = cy[idx[lo]], y1 = y0;
    for (int q = lo + 1; q < hi; ++q) {
      const int e = idx[q];
      x0 = std::min(x0, cx[e]); x1 = std::max(x1, cx[e]);
      y0 = std::min(y0, cy[e]); y1 = std::max(y1, cy[e]);
    }
    const std::vector<double>& c = (x1 - x0 >= y1 - y0) ? cx : cy;
    const int k1 = k / 2;
    int m = lo + (int)((int64_t)(hi - lo) * k1 / k);
    m = std::max(lo + 1, std::min(hi - 1, m));
    std::nth_element(idx.begin() + lo, idx.begin() + m, idx.begin() + hi,
                     [&c](int32_t a, int32_t b) { return c[a] < c[b] || (c[a] == c[b] && a < b); });
    cut(lo, m, k1);
    cut(m, hi, k - k1);
  }
};

template <typename T>
int upload_vec(T** d, const std::vector<T>& v) {
  if (*d) { hipFree(*d); *d = nullptr; }
  BL_HIP(hipMalloc(d, sizeof(T) * std::max<size_t>(v.size(), 1)));
  if (!v.empty()) BL_HIP(hipMemcpy(*d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return BL_SUCCESS;
}

}  // namespace

void free_tiles(Handle* h) {
  void* bufs[] = {h->d_tile_eptr, h->d_tile_nptr, h->d_tile_bounds, h->d_tile_nodes, h->d_tile_ncm, h->d_tile_el,
                  h->d_tile_nd, h->d_rest8, h->d_ens_toff};
  for (void* p : bufs)
    if (p) hipFree(p);
  h->d_tile_eptr = h->d_tile_nptr = h->d_tile_nodes = nullptr;
  h->d_tile_ncm = nullptr;
  h->d_tile_bounds = nullptr;
  h->d_tile_el = nullptr;
  h->d_tile_nd = nullptr;
  h->d_ens_toff = nullptr;
  h->d_rest8 = nullptr;
  h->ntile = h->nentries = 0;
  h->tile_state = 0;
}

int build_tiles(Handle* h) {
  free_tiles(h);
  const int nel = h->sz.nel, nnd = h->sz.nnd;
  std::vector<int32_t> ielnd(4 * (size_t)nel);
  std::vector<double> x(nnd), y(nnd);
  bl_handle_t* hh = static_cast<bl_handle_t*>(h);
  if (bl_download(hh, BL_IELND, ielnd.data(), 4 * (int64_t)nel) ||
      bl_download(hh, BL_NDX, x.data(), nnd) || bl_download(hh, BL_NDY, y.data(), nnd))
    return BL_FAILURE;
  // every node is corner jj of at most one element, for each jj
  std::vector<uint8_t> seen(4 * (size_t)nnd, 0);
  std::vector<int32_t> adj_ptr(nnd + 1, 0);
  for (int e = 0; e < nel; ++e)
    for (int j = 0; j < 4; ++j) {
      const int v = ielnd[4 * (size_t)e + j] - 1;
      if (v < 0 || v >= nnd || seen[4 * (size_t)v + j]) {
        h->tile_state = -1;  // not a tiling this step can use
        return BL_SUCCESS;
      }
      seen[4 * (size_t)v + j] = 1;
      ++adj_ptr[v + 1];
    }
  for (int v = 0; v < nnd; ++v) {
    // a node no element uses would be written by no tile
    if (adj_ptr[v + 1] == 0) {
      h->tile_state = -1;
      return BL_SUCCESS;
    }
    adj_ptr[v + 1] += adj_ptr[v];
  }
  // (an element with a repeated node is listed once per corner: the counts
  // below take each element once)
  std::vector<int32_t> adj(adj_ptr[nnd]), fill(adj_ptr.begin(), adj_ptr.end() - 1);
  for (int e = 0; e < nel; ++e)
    for (int j = 0; j < 4; ++j) adj[fill[ielnd[4 * (size_t)e + j] - 1]++] = e;
  if (nel >= (1 << 30) || nnd >= (1 << 31) - 1) { h->tile_state = -1; return BL_SUCCESS; }
  std::vector<double> cx(nel), cy(nel);
  for (int e = 0; e < nel; ++e) {
    const int32_t* q = &ielnd[4 * (size_t)e];
    cx[e] = 0.25 * (x[q[0] - 1] + x[q[1] - 1] + x[q[2] - 1] + x[q[3] - 1]);
    cy[e] = 0.25 * (y[q[0] - 1] + y[q[1] - 1] + y[q[2] - 1] + y[q[3] - 1]);
  }
  Rcb r{cx, cy, ielnd, adj_ptr, adj, {}, std::vector<int32_t>(nnd, 0),
        std::vector<int32_t>(nel, 0), {}, 0};
  r.idx.resize(nel);
  for (int e = 0; e < nel; ++e) r.idx[e] = e;
  // an ensemble's instances own contiguous element ranges: each is cut on
  // its own, so that a tile lies in one instance (one dt)
  std::vector<int32_t> toff;
  if (h->ens_n > 0) {
    toff.assign(h->ens_n + 1, 0);
    for (int k = 0; k < h->ens_n; ++k) {
      const int lo = h->ens_eoff[k], hi = h->ens_eoff[k + 1];
      toff[k] = (int)r.leaves.size();
      r.cut(lo, hi, std::max(1, (hi - lo + kTileTarget - 1) / kTileTarget));
    }
    toff[h->ens_n] = (int)r.leaves.size();
  } else {
    r.cut(0, nel, std::max(1, (nel + kTileTarget - 1) / kTileTarget));
  }

  const int ntile = (int)r.leaves.size();
  std::vector<int32_t> tile_of(nel);
  for (int t = 0; t < ntile; ++t)
    for (int k = r.leaves[t].first; k < r.leaves[t].second; ++k) tile_of[r.idx[k]] = t;
  // per tile: its nodes (ascending; the high bit marks the tile that writes
  // the node, the one holding the node's first incident element), then the
  // entries: its own elements ascending, then the ring around them
  std::vector<int32_t> nptr(ntile + 1, 0), nodes, eptr(ntile + 1, 0);
  std::vector<uint8_t> ncm;
  std::vector<int2> ents;
  std::vector<int4> ent_nd;
  nodes.reserve((size_t)nnd + nnd / 4);
  ncm.reserve((size_t)nnd + nnd / 4);
  ents.reserve((size_t)nel + nel / 2);
  ent_nd.reserve((size_t)nel + nel / 2);
  std::vector<int32_t> nstamp(nnd, -1), slot_of(nnd, 0), estamp(nel, -1), tn, own, ring;
  for (int t = 0; t < ntile; ++t) {
    own.assign(r.idx.begin() + r.leaves[t].first, r.idx.begin() + r.leaves[t].second);
    std::sort(own.begin(), own.end());
    tn.clear();
    for (int e : own) {
      estamp[e] = t;
      for (int j = 0; j < 4; ++j) {
        const int v = ielnd[4 * (size_t)e + j] - 1;
        if (nstamp[v] != t) { nstamp[v] = t; tn.push_back(v); }
      }
    }
    std::sort(tn.begin(), tn.end());
    ring.clear();
    for (int v : tn)
      for (int q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q)
        if (estamp[adj[q]] != t) { estamp[adj[q]] = t; ring.push_back(adj[q]); }
    std::sort(ring.begin(), ring.end());
    if ((int)tn.size() > kTileNodes || own.size() + ring.size() > (size_t)kTileEls) {
      h->tile_state = -1;
      return BL_SUCCESS;
    }
    nptr[t] = (int)nodes.size();
    for (size_t k = 0; k < tn.size(); ++k) {
      const int v = tn[k];
      slot_of[v] = (int)k;
      const bool writes = tile_of[adj[adj_ptr[v]]] == t;
      nodes.push_back((int32_t)((uint32_t)v | (writes ? 0x80000000u : 0u)));
      // the corners jj the node is (every element around it is an entry
      // of this tile, own or ring)
      uint8_t cm = 0;
      for (int j = 0; j < 4; ++j) cm |= seen[4 * (size_t)v + j] ? (uint8_t)(1u << j) : 0u;
      ncm.push_back(cm);
    }
    eptr[t] = (int)ents.size();
    for (int pass = 0; pass < 2; ++pass)
      for (int e : pass == 0 ? own : ring) {
        uint32_t slots = 0;
        for (int j = 0; j < 4; ++j) {
          const int v = ielnd[4 * (size_t)e + j] - 1;
          slots |= (uint32_t)(nstamp[v] == t ? slot_of[v] : kNoSlot) << (8 * j);
        }
        // an own element: this tile completes and writes it
        const uint32_t flags = pass == 0 ? kTileOwn : 0u;
        ents.push_back(make_int2((int)((uint32_t)e | (flags << 30)), (int)slots));
        const int32_t* q = &ielnd[4 * (size_t)e];
        ent_nd.push_back(make_int4(q[0] - 1, q[1] - 1, q[2] - 1, q[3] - 1));
      }
  }
  nptr[ntile] = (int)nodes.size();
  eptr[ntile] = (int)ents.size();
  std::vector<int4> bounds(ntile);
  for (int t = 0; t < ntile; ++t) bounds[t] = make_int4(eptr[t], eptr[t + 1], nptr[t], nptr[t + 1]);
  if (upload_vec(&h->d_tile_eptr, eptr) || upload_vec(&h->d_tile_nptr, nptr) ||
      upload_vec(&h->d_tile_bounds, bounds) ||
      upload_vec(&h->d_tile_nodes, nodes) || upload_vec(&h->d_tile_ncm, ncm) ||
      upload_vec(&h->d_tile_el, ents) ||
      upload_vec(&h->d_tile_nd, ent_nd) ||
      (!toff.empty() && upload_vec(&h->d_ens_toff, toff)))
    return BL_FAILURE;
  BL_HIP(hipMalloc(&h->d_rest8, 2 * (size_t)std::max(nel, 1)));
  BL_HIP(hipMemset(h->d_rest8, 0, 2 * (size_t)std::max(nel, 1)));
  h->ntile = ntile;
  h->nentries = (int32_t)ents.size();
  // the partial slots of the tile step: one per tile
  if (ntile + 16 > h->npart) {
    hipFree(h->d_part);
    hipFree(h->d_parti);
    h->d_part = nullptr;
    h->d_parti = nullptr;
    h->npart = ntile + 16;
    BL_HIP(hipMalloc(&h->d_part, sizeof(double) * 4 * h->npart));
    BL_HIP(hipMalloc(&h->d_parti, sizeof(int32_t) * 4 * h->npart));
  }
  // the second element set
  if (!h->elset[0][0]) {
    const int ids[4] = {BL_ELDENSITY, BL_ELENERGY, BL_ELPRESSURE, BL_ELCS2};
    for (int k = 0; k < 4; ++k) {
      h->elset[h->cur_set][k] = h->R(ids[k]);
      BL_HIP(hipMalloc(&h->elset[1 - h->cur_set][k],
                       sizeof(double) * std::max<int64_t>(h->count[ids[k]], 1)));
    }
    h->el_cur = h->cur_set;
  }
  h->tile_state = 1;
  return BL_SUCCESS;
}

}  // namespace bl
