// Host-side preparation of the solver's device layout (see host_prep.h).
#include "host_prep.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "mtcocr.h"

namespace mtcocr {

namespace {

struct Entry {
  int col;
  double a, b;
  bool cplx;
};

// Row lists of the full matrix, built from lower-triangle blocks: row i receives its lower
// blocks (ascending), its diagonal, then the transposed blocks of rows k > i (ascending k), so
// columns are ascending within every row (the order of FoldedPreconditioner::From2x2ToCSR2x2_2,
// Code/COCR_FP/FoldedPreconditioner.cpp:653-736).
template <class GetLower, class GetDiag>
int build_rows(int nb, const int* ig, const int* jg, GetLower lower, GetDiag diag,
               const int* order, SlicedEll& out) {
  if (nb <= 0 || !ig || !jg) return MTCOCR_ERR_ARG;
  const int64_t nnzl = ig[nb];
  std::vector<int64_t> cnt(nb + 1, 0);
  for (int i = 0; i < nb; ++i) {
    cnt[i] += 1 + (ig[i + 1] - ig[i]);
    for (int64_t k = ig[i]; k < ig[i + 1]; ++k) {
      const int j = jg[k];
      if (j < 0 || j >= i) return MTCOCR_ERR_ARG;
      cnt[j] += 1;
    }
  }
  std::vector<int64_t> ptr(nb + 1, 0);
  for (int i = 0; i < nb; ++i) ptr[i + 1] = ptr[i] + cnt[i];
  std::vector<Entry> ent(ptr[nb]);
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  out.diag_a.assign(nb, 0.0);
  out.diag_b.assign(nb, 0.0);
  for (int i = 0; i < nb; ++i) {
    for (int64_t k = ig[i]; k < ig[i + 1]; ++k) {
      Entry e = lower(k);
      e.col = jg[k];
      ent[fill[i]++] = e;
    }
    Entry d = diag(i);
    d.col = i;
    out.diag_a[i] = d.a;
    out.diag_b[i] = d.b;
    ent[fill[i]++] = d;
  }
  for (int i = 0; i < nb; ++i)
    for (int64_t k = ig[i]; k < ig[i + 1]; ++k) {
      Entry e = lower(k);
      e.col = i;
      ent[fill[jg[k]]++] = e;
    }
  (void)nnzl;
  // Device row i holds the reference's row order[i] (identity without an order); columns are
  // renumbered through the inverse and every row keeps the reference's entry order, so each
  // row's sum is formed in the same order whatever the permutation.
  std::vector<int> inv;
  if (order) {
    inv.assign(nb, -1);
    for (int i = 0; i < nb; ++i) {
      const int o = order[i];
      if (o < 0 || o >= nb || inv[o] != -1) return MTCOCR_ERR_ARG;
      inv[o] = i;
    }
  }
  auto old_of = [&](int i) { return order ? order[i] : i; };
  auto new_col = [&](int c) { return order ? inv[c] : c; };
  {
    std::vector<double> da(nb), db(nb);
    for (int i = 0; i < nb; ++i) {
      da[i] = out.diag_a[old_of(i)];
      db[i] = out.diag_b[old_of(i)];
    }
    out.diag_a.swap(da);
    out.diag_b.swap(db);
  }
  // slices of 64 rows: per slice the maximum complex / real counts (explicit zeros dropped,
  // the diagonal always kept)
  const int ns = (nb + 63) / 64;
  out.nb = nb;
  out.nslices = ns;
  out.offc.assign(ns + 1, 0);
  out.offr.assign(ns + 1, 0);
  out.nc.assign(ns, 0);
  out.nr.assign(ns, 0);
  std::vector<int> rc(nb, 0), rr(nb, 0);
  std::vector<char> diag_cplx(nb, 0);
  out.nnz_complex = out.nnz_real = 0;
  for (int i = 0; i < nb; ++i) {
    const int oi = old_of(i);
    for (int64_t k = ptr[oi]; k < ptr[oi + 1]; ++k) {
      const Entry& e = ent[k];
      if (e.col == oi) diag_cplx[i] = e.cplx;
      const bool keep = (e.col == oi) || e.a != 0.0 || e.b != 0.0;
      if (!keep) continue;
      if (e.cplx) ++rc[i]; else ++rr[i];
    }
    out.nnz_complex += rc[i];
    out.nnz_real += rr[i];
    const int s = i >> 6;
    out.nc[s] = std::max(out.nc[s], rc[i]);
    out.nr[s] = std::max(out.nr[s], rr[i]);
  }
  for (int s = 0; s < ns; ++s) {
    const int64_t oc = (int64_t)out.offc[s] + 64LL * out.nc[s];
    const int64_t orr = (int64_t)out.offr[s] + 64LL * out.nr[s];
    if (oc > 2147483000LL || orr > 2147483000LL) return MTCOCR_ERR_ARG;
    out.offc[s + 1] = int(oc);
    out.offr[s + 1] = int(orr);
  }
  if (out.nnz_complex + out.nnz_real > 2147483000LL) return MTCOCR_ERR_ARG;
  // row pack: the diagonal (always stored, complex or real) is left out
  out.rp.assign(nb + 1, 0);
  out.rb.assign(nb + 1, 0);
  for (int i = 0; i < nb; ++i) {
    const int offc = rc[i] - (diag_cplx[i] ? 1 : 0), offr = rr[i] - (diag_cplx[i] ? 0 : 1);
    out.rp[i + 1] = out.rp[i] + offc + offr;
    out.rb[i + 1] = out.rb[i] + offc;
  }
  out.pcol.assign(out.rp[nb], 0);
  out.pa.assign(out.rp[nb], 0.0);
  out.pb.assign(out.rb[nb], 0.0);
  const int64_t totc = out.offc[ns], totr = out.offr[ns];
  out.colc.assign(totc, 0);
  out.ac.assign(totc, 0.0);
  out.bc.assign(totc, 0.0);
  out.colr.assign(totr, 0);
  out.ar.assign(totr, 0.0);
  for (int i = 0; i < nb; ++i) {
    const int oi = old_of(i);
    const int s = i >> 6, lane = i & 63;
    int kc = 0, kr = 0;
    for (int64_t k = ptr[oi]; k < ptr[oi + 1]; ++k) {
      const Entry& e = ent[k];
      const bool keep = (e.col == oi) || e.a != 0.0 || e.b != 0.0;
      if (!keep) continue;
      if (e.cplx) {
        const int64_t p = out.offc[s] + 64LL * kc + lane;
        out.colc[p] = new_col(e.col);
        out.ac[p] = e.a;
        out.bc[p] = e.b;
        ++kc;
      } else {
        const int64_t p = out.offr[s] + 64LL * kr + lane;
        out.colr[p] = new_col(e.col);
        out.ar[p] = e.a;
        ++kr;
      }
    }
    // row pack: complex off-diagonals, then real off-diagonals (each in the row's order)
    int64_t q = out.rp[i], qb = out.rb[i];
    for (int pass = 0; pass < 2; ++pass)
      for (int64_t k = ptr[oi]; k < ptr[oi + 1]; ++k) {
        const Entry& e = ent[k];
        if (e.col == oi || (e.a == 0.0 && e.b == 0.0) || e.cplx != (pass == 0)) continue;
        out.pcol[q] = new_col(e.col);
        out.pa[q++] = e.a;
        if (pass == 0) out.pb[qb++] = e.b;
      }
    // padding slots point at the row itself with zero value
    for (; kc < out.nc[s]; ++kc) out.colc[out.offc[s] + 64LL * kc + lane] = i;
    for (; kr < out.nr[s]; ++kr) out.colr[out.offr[s] + 64LL * kr + lane] = i;
  }
  // lanes past nb in the last slice keep column 0 and zero values
  // row blocks of kBlockRows rows and their halos
  {
    constexpr int R = SlicedEll::kBlockRows;
    const int nblk = (nb + R - 1) / R;
    out.hoff.assign(nblk + 1, 0);
    out.halo.clear();
    out.halo.reserve((size_t)nb * 10);
    out.lcol.assign(out.pcol.size(), 0);
    out.ldiag.assign(nb, 0);
    out.halo_max = 0;
    std::vector<int> stamp(nb, -1), pos(nb, 0), cols;
    for (int k = 0; k < nblk && out.halo_max <= SlicedEll::kHaloCap; ++k) {
      const int r0 = k * R, r1 = std::min(nb, r0 + R);
      cols.clear();
      auto take = [&](int c) {
        if (stamp[c] != k) {
          stamp[c] = k;
          cols.push_back(c);
        }
      };
      for (int q = r0; q < r1; ++q) {
        take(q);
        for (int e = out.rp[q]; e < out.rp[q + 1]; ++e) take(out.pcol[e]);
      }
      std::sort(cols.begin(), cols.end());
      for (size_t i = 0; i < cols.size(); ++i) pos[cols[i]] = (int)i;
      for (int q = r0; q < r1; ++q) {
        out.ldiag[q] = (unsigned short)pos[q];
        for (int e = out.rp[q]; e < out.rp[q + 1]; ++e)
          out.lcol[e] = (unsigned short)pos[out.pcol[e]];
      }
      out.halo.insert(out.halo.end(), cols.begin(), cols.end());
      out.hoff[k + 1] = (int)out.halo.size();
      out.halo_max = std::max(out.halo_max, (int)cols.size());
    }
    if (out.halo_max > SlicedEll::kHaloCap) {
      out.hoff.clear();
      out.halo.clear();
      out.lcol.clear();
      out.ldiag.clear();
    }
  }
  return 0;
}

}  // namespace

int ell_from_block2x2(int nb, const int* ig, const int* jg, const int* idi, const int* ijg,
                      const double* di, const double* gg, const int* order, SlicedEll& out) {
  if (!idi || !ijg || !di || !gg) return MTCOCR_ERR_ARG;
  auto lower = [&](int64_t k) {
    Entry e;
    const int beg = ijg[k], sz = ijg[k + 1] - ijg[k];
    e.a = gg[beg];
    e.b = (sz == 2) ? gg[beg + 1] : 0.0;
    e.cplx = (e.b != 0.0);
    e.col = 0;
    return e;
  };
  auto diag = [&](int i) {
    Entry e;
    const int beg = idi[i], sz = idi[i + 1] - idi[i];
    e.a = di[beg];
    e.b = (sz == 2) ? di[beg + 1] : 0.0;
    e.cplx = (e.b != 0.0);
    e.col = i;
    return e;
  };
  return build_rows(nb, ig, jg, lower, diag, order, out);
}

int ell_from_ks(int nb, const int* ig, const int* jg, const double* k_di, const double* k_gg,
                const double* s_di, const double* s_gg, const int* order, SlicedEll& out) {
  if (!k_di || !k_gg || !s_di || !s_gg) return MTCOCR_ERR_ARG;
  auto lower = [&](int64_t k) {
    Entry e;
    e.a = k_gg[k];
    e.b = s_gg[k];
    e.cplx = (e.b != 0.0);
    e.col = 0;
    return e;
  };
  auto diag = [&](int i) {
    Entry e;
    e.a = k_di[i];
    e.b = s_di[i];
    e.cplx = (e.b != 0.0);
    e.col = i;
    return e;
  };
  return build_rows(nb, ig, jg, lower, diag, order, out);
}

int gradient_csr(int kpar, int n_edges, int n_nodes_c, const int* nvkat, const int* nver,
                 const double* sigma3d, const int* edges, const int* ig_t, const int* jg_t,
                 const double* gg_t, const int* is_node_bound, GradientCsr& out) {
  if (kpar <= 0 || n_edges <= 0 || n_nodes_c <= 0 || !nvkat || !nver || !sigma3d || !edges ||
      !is_node_bound)
    return MTCOCR_ERR_ARG;
  // nodes of conductive cells, boundary nodes excluded, renumbered in ascending order
  std::vector<char> in_sigma(n_nodes_c, 0);
  for (int i = 0; i < kpar; ++i) {
    if (!(sigma3d[nvkat[i]] > 0.0)) continue;
    for (int j = 0; j < 8; ++j) {
      const int node = nver[(int64_t)i * 14 + j];
      if (node < 0) return MTCOCR_ERR_ARG;
      if (node < n_nodes_c) {
        if (is_node_bound[node]) continue;
        in_sigma[node] = 1;
      }
    }
  }
  std::vector<int> renum(n_nodes_c, -1);
  int m = 0;
  for (int i = 0; i < n_nodes_c; ++i)
    if (in_sigma[i]) renum[i] = m++;
  // G^T rows (edge -> nodes): -1/2 at the first node, +1/2 at the second; hanging nodes
  // (index >= n_nodes_c) spread through the T-matrix ig_t/jg_t/gg_t onto conforming nodes
  out.m = m;
  out.tptr.assign(n_edges + 1, 0);
  out.tcol.clear();
  out.tval.clear();
  out.tcol.reserve((size_t)n_edges * 2);
  out.tval.reserve((size_t)n_edges * 2);
  for (int i = 0; i < n_edges; ++i) {
    for (int k = 0; k < 2; ++k) {
      const int v = edges[(int64_t)i * 2 + k];
      const double sgn = (k == 0) ? -0.5 : 0.5;
      if (v < 0) return MTCOCR_ERR_ARG;
      if (v < n_nodes_c) {
        if (in_sigma[v]) {
          out.tcol.push_back(renum[v]);
          out.tval.push_back(sgn);
        }
      } else {
        if (!ig_t || !jg_t || !gg_t) return MTCOCR_ERR_ARG;
        for (int ii = ig_t[v]; ii < ig_t[v + 1]; ++ii) {
          const int kk = jg_t[ii];
          if (kk < 0 || kk >= n_nodes_c) return MTCOCR_ERR_ARG;  // T rows hold conforming nodes
          if (in_sigma[kk]) {
            out.tcol.push_back(renum[kk]);
            out.tval.push_back(sgn * gg_t[ii]);
          }
        }
      }
    }
    out.tptr[i + 1] = (int64_t)out.tcol.size();
  }
  // G rows (node -> edges, ascending edge): the transpose
  std::vector<int64_t> gcnt(m + 1, 0);
  for (size_t k = 0; k < out.tcol.size(); ++k) gcnt[out.tcol[k]]++;
  out.gptr.assign(m + 1, 0);
  for (int i = 0; i < m; ++i) out.gptr[i + 1] = out.gptr[i] + gcnt[i];
  out.gcol.assign(out.tcol.size(), 0);
  out.gval.assign(out.tcol.size(), 0.0);
  std::vector<int64_t> fill(out.gptr.begin(), out.gptr.end() - 1);
  for (int e = 0; e < n_edges; ++e)
    for (int64_t k = out.tptr[e]; k < out.tptr[e + 1]; ++k) {
      const int node = out.tcol[k];
      out.gcol[fill[node]] = e;
      out.gval[fill[node]] = out.tval[k];
      fill[node]++;
    }
  return 0;
}

int build_gradient(int kpar, int n_edges, int n_nodes_c, const int* nvkat, const int* nver,
                   const double* sigma3d, const int* edges, const int* ig_t, const int* jg_t,
                   const double* gg_t, const int* is_node_bound, const int* edge_order,
                   GradientEll& out) {
  GradientCsr C;
  const int rc = gradient_csr(kpar, n_edges, n_nodes_c, nvkat, nver, sigma3d, edges, ig_t, jg_t,
                              gg_t, is_node_bound, C);
  if (rc) return rc;
  const int m = C.m;
  const std::vector<int64_t>&tptr = C.tptr, &gptr = C.gptr;
  const std::vector<int>&tcol = C.tcol, &gc = C.gcol;
  const std::vector<double>&tval = C.tval, &gv = C.gval;
  // Device order: edges as the matrix rows (edge_order, new -> old); interior nodes ranked by
  // their first incident edge in that order, so that a node's edges and the node sit close.
  // Entry order inside every row of G and G^T stays the reference's.
  std::vector<int> einv, norder(m), ninv(m);
  for (int i = 0; i < m; ++i) norder[i] = ninv[i] = i;
  if (edge_order) {
    einv.assign(n_edges, -1);
    for (int e = 0; e < n_edges; ++e) {
      const int o = edge_order[e];
      if (o < 0 || o >= n_edges || einv[o] != -1) return MTCOCR_ERR_ARG;
      einv[o] = e;
    }
    std::vector<int> first(m, n_edges);
    for (int i = 0; i < m; ++i)
      for (int64_t q = gptr[i]; q < gptr[i + 1]; ++q) first[i] = std::min(first[i], einv[gc[q]]);
    std::stable_sort(norder.begin(), norder.end(),
                     [&](int a, int b) { return first[a] < first[b]; });
    for (int i = 0; i < m; ++i) ninv[norder[i]] = i;
  }
  auto new_edge = [&](int e) { return edge_order ? einv[e] : e; };
  auto old_edge = [&](int e) { return edge_order ? edge_order[e] : e; };
  out.n = n_edges;
  out.m = m;
  out.wg = 0;
  out.wgt = 0;
  for (int i = 0; i < m; ++i) out.wg = std::max<int>(out.wg, int(gptr[i + 1] - gptr[i]));
  for (int e = 0; e < n_edges; ++e) out.wgt = std::max<int>(out.wgt, int(tptr[e + 1] - tptr[e]));
  out.gcol.assign((size_t)out.wg * m, 0);
  out.gval.assign((size_t)out.wg * m, 0.0);
  for (int i = 0; i < m; ++i) {
    const int oi = norder[i];
    int k = 0;
    for (int64_t q = gptr[oi]; q < gptr[oi + 1]; ++q, ++k) {
      out.gcol[(size_t)k * m + i] = new_edge(gc[q]);
      out.gval[(size_t)k * m + i] = gv[q];
    }
  }
  out.gtcol.assign((size_t)out.wgt * n_edges, 0);
  out.gtval.assign((size_t)out.wgt * n_edges, 0.0);
  for (int e = 0; e < n_edges; ++e) {
    const int oe = old_edge(e);
    int k = 0;
    for (int64_t q = tptr[oe]; q < tptr[oe + 1]; ++q, ++k) {
      out.gtcol[(size_t)k * n_edges + e] = ninv[tcol[q]];
      out.gtval[(size_t)k * n_edges + e] = tval[q];
    }
  }
  return 0;
}

bool gradient_on_grid(const GradientEll& G, const GridDims& d) {
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  if (nx < 3 || ny < 3 || nz < 3) return false;
  const int64_t m = (int64_t)(nx - 2) * (ny - 2) * (nz - 2);
  if (G.m != m || G.n != d.base(2) + d.count(2) || G.wg != 6 || G.wgt != 2) return false;
  auto row = [&](int dir, int i, int j, int k) -> int64_t {
    return d.base(dir) + ((int64_t)k * d.nyd(dir) + j) * d.nxd(dir) + i;
  };
  auto node = [&](int i, int j, int k) -> int64_t {
    if (i < 1 || j < 1 || k < 1 || i > nx - 2 || j > ny - 2 || k > nz - 2) return -1;
    return (i - 1) + (int64_t)(nx - 2) * ((j - 1) + (int64_t)(ny - 2) * (k - 1));
  };
  for (int k = 1; k <= nz - 2; ++k)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i) {
        const int64_t n = node(i, j, k);
        const int64_t e[6] = {row(2, i, j, k - 1), row(1, i, j - 1, k), row(0, i - 1, j, k),
                              row(0, i, j, k),     row(1, i, j, k),     row(2, i, j, k)};
        for (int q = 0; q < 6; ++q) {
          if (G.gcol[(size_t)q * m + n] != e[q]) return false;
          if (G.gval[(size_t)q * m + n] != (q < 3 ? 0.5 : -0.5)) return false;
        }
      }
  const int64_t nb = G.n;
  for (int dir = 0; dir < 3; ++dir)
    for (int k = 0; k < d.nzd(dir); ++k)
      for (int j = 0; j < d.nyd(dir); ++j)
        for (int i = 0; i < d.nxd(dir); ++i) {
          const int64_t e = row(dir, i, j, k);
          const int64_t lo = node(i, j, k);
          const int64_t up = node(i + (dir == 0), j + (dir == 1), k + (dir == 2));
          int64_t want_c[2] = {0, 0};
          double want_v[2] = {0.0, 0.0};
          int c = 0;
          if (lo >= 0) {
            want_c[c] = lo;
            want_v[c++] = -0.5;
          }
          if (up >= 0) {
            want_c[c] = up;
            want_v[c++] = 0.5;
          }
          for (int q = 0; q < 2; ++q) {
            if (G.gtval[(size_t)q * nb + e] != want_v[q]) return false;
            if (q < c && G.gtcol[(size_t)q * nb + e] != want_c[q]) return false;
          }
        }
  return true;
}

namespace {

constexpr int kUnset = std::numeric_limits<int>::min();

// Integer node coordinates of a structured hexahedral mesh from the cells' local vertex order
// (vertex l of a cell sits at offset (l&1, (l>>1)&1, (l>>2)&1) from vertex 0), propagated cell
// to cell; nodes of no cell stay kUnset. Fails when the cells are not structured hexahedra.
int mesh_coords(int kpar, int n_nodes, const int* nver, std::vector<int>& cx,
                std::vector<int>& cy, std::vector<int>& cz) {
  cx.assign(n_nodes, kUnset);
  cy.assign(n_nodes, kUnset);
  cz.assign(n_nodes, kUnset);
  std::vector<int64_t> nptr(n_nodes + 1, 0);
  for (int64_t c = 0; c < kpar; ++c)
    for (int l = 0; l < 8; ++l) {
      const int v = nver[c * 14 + l];
      if (v < 0 || v >= n_nodes) return MTCOCR_ERR_ARG;
      nptr[v + 1]++;
    }
  for (int v = 0; v < n_nodes; ++v) nptr[v + 1] += nptr[v];
  std::vector<int> ncell(nptr[n_nodes]);
  {
    std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
    for (int c = 0; c < kpar; ++c)
      for (int l = 0; l < 8; ++l) ncell[fill[nver[(int64_t)c * 14 + l]]++] = c;
  }
  std::vector<char> done(kpar, 0);
  std::vector<int> queue;
  queue.reserve(kpar);
  for (int seed = 0; seed < kpar; ++seed) {
    if (done[seed]) continue;
    const int v0 = nver[(int64_t)seed * 14];
    if (cx[v0] == kUnset) cx[v0] = cy[v0] = cz[v0] = 0;
    done[seed] = 1;
    queue.push_back(seed);
    for (size_t head = queue.size() - 1; head < queue.size(); ++head) {
      const int c = queue[head];
      const int* v = nver + (int64_t)c * 14;
      int bx = kUnset, by = 0, bz = 0;
      for (int l = 0; l < 8 && bx == kUnset; ++l)
        if (cx[v[l]] != kUnset) {
          bx = cx[v[l]] - (l & 1);
          by = cy[v[l]] - ((l >> 1) & 1);
          bz = cz[v[l]] - ((l >> 2) & 1);
        }
      if (bx == kUnset) return MTCOCR_ERR_ARG;
      for (int l = 0; l < 8; ++l) {
        const int x = bx + (l & 1), y = by + ((l >> 1) & 1), z = bz + ((l >> 2) & 1);
        if (cx[v[l]] == kUnset) {
          cx[v[l]] = x;
          cy[v[l]] = y;
          cz[v[l]] = z;
        } else if (cx[v[l]] != x || cy[v[l]] != y || cz[v[l]] != z) {
          return MTCOCR_ERR_ARG;  // not a structured hexahedral mesh in this vertex order
        }
        for (int64_t q = nptr[v[l]]; q < nptr[v[l] + 1]; ++q)
          if (!done[ncell[q]]) {
            done[ncell[q]] = 1;
            queue.push_back(ncell[q]);
          }
      }
    }
  }
  return 0;
}

}  // namespace

int order_from_mesh(int kpar, int n_nodes, const int* nver, int n_edges, const int* edges,
                    int block, std::vector<int>& order) {
  if (kpar <= 0 || n_nodes <= 0 || n_edges <= 0 || !nver || !edges || block < 1)
    return MTCOCR_ERR_ARG;
  std::vector<int> cx, cy, cz;
  const int rc = mesh_coords(kpar, n_nodes, nver, cx, cy, cz);
  if (rc) return rc;
  // edge key: (pencil x, pencil y, z, y, x, direction) of the edge's lower end
  struct Key {
    int px, py, z, y, x, d;
  };
  std::vector<Key> key(n_edges);
  for (int e = 0; e < n_edges; ++e) {
    const int a = edges[(int64_t)e * 2], b = edges[(int64_t)e * 2 + 1];
    if (a < 0 || b < 0 || a >= n_nodes || b >= n_nodes || cx[a] == kUnset || cx[b] == kUnset)
      return MTCOCR_ERR_ARG;
    const int x = std::min(cx[a], cx[b]), y = std::min(cy[a], cy[b]), z = std::min(cz[a], cz[b]);
    const int d = cx[a] != cx[b] ? 0 : (cy[a] != cy[b] ? 1 : 2);
    key[e] = Key{x >= 0 ? x / block : -1 - (-x - 1) / block, y >= 0 ? y / block : -1 - (-y - 1) / block,
                 z, y, x, d};
  }
  order.resize(n_edges);
  for (int e = 0; e < n_edges; ++e) order[e] = e;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const Key &p = key[a], &q = key[b];
    if (p.px != q.px) return p.px < q.px;
    if (p.py != q.py) return p.py < q.py;
    if (p.z != q.z) return p.z < q.z;
    if (p.y != q.y) return p.y < q.y;
    if (p.x != q.x) return p.x < q.x;
    return p.d < q.d;
  });
  return 0;
}

int grid_order_from_mesh(int kpar, int n_nodes, const int* nver, int n_edges, const int* edges,
                         std::vector<int>& order, GridDims& g) {
  if (kpar <= 0 || n_nodes <= 0 || n_edges <= 0 || !nver || !edges) return MTCOCR_ERR_ARG;
  std::vector<int> cx, cy, cz;
  int rc = mesh_coords(kpar, n_nodes, nver, cx, cy, cz);
  if (rc) return rc;
  int lo[3] = {std::numeric_limits<int>::max(), std::numeric_limits<int>::max(),
               std::numeric_limits<int>::max()};
  int hi[3] = {kUnset, kUnset, kUnset};
  for (int v = 0; v < n_nodes; ++v) {
    if (cx[v] == kUnset) return MTCOCR_ERR_ARG;  // a node of no cell
    const int c[3] = {cx[v], cy[v], cz[v]};
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], c[a]);
      hi[a] = std::max(hi[a], c[a]);
    }
  }
  g.nx = hi[0] - lo[0] + 1;
  g.ny = hi[1] - lo[1] + 1;
  g.nz = hi[2] - lo[2] + 1;
  if ((int64_t)g.nx * g.ny * g.nz != n_nodes) return MTCOCR_ERR_ARG;
  if (g.count(0) + g.count(1) + g.count(2) != n_edges) return MTCOCR_ERR_ARG;
  order.assign(n_edges, -1);
  for (int e = 0; e < n_edges; ++e) {
    const int a = edges[(int64_t)e * 2], b = edges[(int64_t)e * 2 + 1];
    if (a < 0 || b < 0 || a >= n_nodes || b >= n_nodes) return MTCOCR_ERR_ARG;
    const int dx = cx[b] - cx[a], dy = cy[b] - cy[a], dz = cz[b] - cz[a];
    if (std::abs(dx) + std::abs(dy) + std::abs(dz) != 1) return MTCOCR_ERR_ARG;
    const int d = dx ? 0 : (dy ? 1 : 2);
    const int i = std::min(cx[a], cx[b]) - lo[0], j = std::min(cy[a], cy[b]) - lo[1],
              k = std::min(cz[a], cz[b]) - lo[2];
    const int64_t r = g.base(d) + ((int64_t)k * g.nyd(d) + j) * g.nxd(d) + i;
    if (order[r] != -1) return MTCOCR_ERR_ARG;  // two edges on one slot
    order[r] = e;
  }
  return 0;
}

int stencil_from_rows(const SlicedEll& E, const GridDims& g, std::vector<double>& coef) {
  const int nb = E.nb;
  if (nb != g.count(0) + g.count(1) + g.count(2) || E.rp.size() != (size_t)nb + 1)
    return MTCOCR_ERR_ARG;
  auto decode = [&](int r, int& d, int& i, int& j, int& k) {
    d = r < g.base(1) ? 0 : (r < g.base(2) ? 1 : 2);
    int64_t q = r - g.base(d);
    i = int(q % g.nxd(d));
    q /= g.nxd(d);
    j = int(q % g.nyd(d));
    k = int(q / g.nyd(d));
  };
  // slot of (column direction, di, dj, dk) within a type-d row, -1 outside the stencil
  auto slot = [&](int d, int dt, int di, int dj, int dk) {
    for (int c = 0; c < kStcCols[d]; ++c) {
      const StencilCol& s = kStc[d][c];
      if (s.dt == dt && s.di == di && s.dj == dj && dk >= s.lo && dk <= s.hi)
        return stc_slot(d, c, dk);
    }
    return -1;
  };
  coef.assign((size_t)nb * kStcCoef, 0.0);
  for (int r = 0; r < nb; ++r) {
    int d, i, j, k;
    decode(r, d, i, j, k);
    double* row = coef.data() + (size_t)r * kStcCoef;
    const int sd = slot(d, d, 0, 0, 0);
    row[sd] = E.diag_a[r];
    row[sd + 1] = E.diag_b[r];
    const int ncx = E.rb[r + 1] - E.rb[r];
    for (int q = E.rp[r]; q < E.rp[r + 1]; ++q) {
      int d2, i2, j2, k2;
      decode(E.pcol[q], d2, i2, j2, k2);
      const int s = slot(d, d2, i2 - i, j2 - j, k2 - k);
      if (s < 0) return MTCOCR_ERR_ARG;
      const bool cplx = (q - E.rp[r]) < ncx;  // the row pack holds complex entries first
      if (d2 == d) {
        row[s] = E.pa[q];
        row[s + 1] = cplx ? E.pb[E.rb[r] + (q - E.rp[r])] : 0.0;
      } else {
        if (cplx) return MTCOCR_ERR_ARG;
        row[s] = E.pa[q];
      }
    }
  }
  return 0;
}

}  // namespace mtcocr

extern "C" int mtcocr_mesh_order(int kpar, int n_nodes, const int* nver, int n_edges,
                                 const int* edges, int block, int* order) {
  if (!order) return MTCOCR_ERR_ARG;
  std::vector<int> ord;
  mtcocr::GridDims dims;
  const int rc =
      block == MTCOCR_ORDER_GRID
          ? mtcocr::grid_order_from_mesh(kpar, n_nodes, nver, n_edges, edges, ord, dims)
          : mtcocr::order_from_mesh(kpar, n_nodes, nver, n_edges, edges, block, ord);
  if (rc) return rc;
  std::memcpy(order, ord.data(), sizeof(int) * (size_t)n_edges);
  return 0;
}

// FoldedPreconditioner::BuildGMatrix (Code/COCR_FP/FoldedPreconditioner.cpp:68-259) on the host:
// the same CSR arrays (0-based), G^T first, in the reference's node numbering and entry order.
extern "C" int mtcocr_build_gmatrix(int kpar, int n_edges_c, int n_nodes_c, const int* nvkat,
                                    const int* nver, const double* sigma3d, const int* edges,
                                    const int* ig_t, const int* jg_t, const double* gg_t,
                                    const int* is_node_bound, int* m, int64_t* nnz,
                                    int* iptr_gt, int* jptr_gt, double* gg_gt, int* iptr_g,
                                    int* jptr_g, double* gg_g) {
  if (!m || !nnz) return MTCOCR_ERR_ARG;
  mtcocr::GradientCsr C;
  const int rc = mtcocr::gradient_csr(kpar, n_edges_c, n_nodes_c, nvkat, nver, sigma3d, edges,
                                      ig_t, jg_t, gg_t, is_node_bound, C);
  if (rc) return rc;
  *m = C.m;
  *nnz = (int64_t)C.tcol.size();
  if (*nnz > 2147483647LL) return MTCOCR_ERR_ARG;
  const bool fill_t = iptr_gt && jptr_gt && gg_gt, fill_g = iptr_g && jptr_g && gg_g;
  if (fill_t) {
    for (int e = 0; e <= n_edges_c; ++e) iptr_gt[e] = int(C.tptr[e]);
    std::memcpy(jptr_gt, C.tcol.data(), sizeof(int) * C.tcol.size());
    std::memcpy(gg_gt, C.tval.data(), sizeof(double) * C.tval.size());
  }
  if (fill_g) {
    for (int i = 0; i <= C.m; ++i) iptr_g[i] = int(C.gptr[i]);
    std::memcpy(jptr_g, C.gcol.data(), sizeof(int) * C.gcol.size());
    std::memcpy(gg_g, C.gval.data(), sizeof(double) * C.gval.size());
  }
  return 0;
}
