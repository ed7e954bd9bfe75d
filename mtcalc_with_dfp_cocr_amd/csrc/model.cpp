// Regular-mesh model layer: topology, portrait, element matrices, 1-D normal field and
// assembly of the complex-symmetric edge-element system A = K/mu + i*omega*sigma*M.
//
// This is the producer of the COCR solver's input. It follows, for regular tensor meshes:
//   numbering / topology  Code/RegularMeshBuilder/main.cpp:923-1031 (BuildEdges,
//                         BuildElementsByEdges), :1389-1447 (WriteNver), :1706-1735 (WriteL13D)
//   portrait              Code/BuildMatrix/T_Portrait.cpp:47-188, T_Portrait.h:37-51
//   element matrices      Code/BuildMatrix/T_Brick.cpp:123-455 (B: curl-curl, C: mass)
//   element RHS           Code/BuildMatrix/T_Brick.cpp:1361-1495
//   global assembly       Code/BuildMatrix/t_global_slae.cpp:80-172
//   Dirichlet conditions  Code/BuildMatrix/bound_cond_vec_harm.cpp:81-108,260-307
//   1-D normal field      Code/Harm1D/mesh_1d_mtz.cpp:95-187, local_matrix_1d.cpp:45-103,
//                         global_slae_1d_harm_prof.cpp:74-138, divgrad_1d.cpp:147-200
// The element matrices are computed from an exact tensor-product integration of the scaled
// lowest-order edge basis (edge DOF = tangential field * h/2), not from the reference's
// expanded closed forms; the two agree to rounding.
#include "mtmodel.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;

struct Grid {
  long long nxn, nyn, nzn;
  long long nxy() const { return nxn * nyn; }
  long long nodes() const { return nxn * nyn * nzn; }
  long long cells() const { return (nxn - 1) * (nyn - 1) * (nzn - 1); }
  long long edges() const {
    return (nxn - 1) * nyn * nzn + nxn * (nyn - 1) * nzn + nxn * nyn * (nzn - 1);
  }
  long long node(long long ix, long long iy, long long iz) const {
    return (iz * nyn + iy) * nxn + ix;
  }
};

bool grid_ok(int nxn, int nyn, int nzn) { return nxn >= 2 && nyn >= 2 && nzn >= 2; }

// Edge numbering: for every node in order, its +x, +y, +z edges (when they exist). This is
// the std::set<pair<int,int>> order of (node, node+1) < (node, node+nx) < (node, node+nxy).
struct EdgeNumbering {
  Grid g;
  explicit EdgeNumbering(const Grid& gg) : g(gg) {}
  // index of the first edge whose first node is (ix,iy,iz): complete z-planes below iz, complete
  // y-rows below iy inside plane iz, then the nodes before ix in row iy
  long long first_edge(long long ix, long long iy, long long iz) const {
    const bool zt = (iz < g.nzn - 1), yt = (iy < g.nyn - 1);
    const long long per_plane = (g.nxn - 1) * g.nyn + g.nxn * (g.nyn - 1) + g.nxn * g.nyn;
    const long long per_row = (g.nxn - 1) + g.nxn + (zt ? g.nxn : 0);
    const long long per_node = 1 + (yt ? 1 : 0) + (zt ? 1 : 0);
    return iz * per_plane + iy * per_row + ix * per_node;
  }
  long long xedge(long long ix, long long iy, long long iz) const { return first_edge(ix, iy, iz); }
  long long yedge(long long ix, long long iy, long long iz) const {
    return first_edge(ix, iy, iz) + (ix < g.nxn - 1 ? 1 : 0);
  }
  long long zedge(long long ix, long long iy, long long iz) const {
    return first_edge(ix, iy, iz) + (ix < g.nxn - 1 ? 1 : 0) + (iy < g.nyn - 1 ? 1 : 0);
  }
  // the 12 edges of cell (cx,cy,cz) in the reference's local order
  void cell_edges(long long cx, long long cy, long long cz, long long* e) const {
    e[0] = xedge(cx, cy, cz);
    e[1] = xedge(cx, cy + 1, cz);
    e[2] = xedge(cx, cy, cz + 1);
    e[3] = xedge(cx, cy + 1, cz + 1);
    e[4] = yedge(cx, cy, cz);
    e[5] = yedge(cx, cy, cz + 1);
    e[6] = yedge(cx + 1, cy, cz);
    e[7] = yedge(cx + 1, cy, cz + 1);
    e[8] = zedge(cx, cy, cz);
    e[9] = zedge(cx + 1, cy, cz);
    e[10] = zedge(cx, cy + 1, cz);
    e[11] = zedge(cx + 1, cy + 1, cz);
  }
};

// Direction (0 x, 1 y, 2 z) and the position signs of the 12 local edges on [-1,1]^3:
// edge i of direction d sits at sign[k] in the two other coordinates k.
constexpr int kLocalDir[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
constexpr int kLocalSign[12][3] = {
    {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},    // x-edges: (y, z)
    {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {1, 0, 1},    // y-edges: (x, z)
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};   // z-edges: (x, y)

// A product  coef * prod_k (p0[k] + p1[k]*xi_k)  on the reference cube.
struct Term {
  double coef;
  double p0[3], p1[3];
};

// integral over [-1,1] of (a0 + a1 t)(b0 + b1 t)
inline double int1(double a0, double a1, double b0, double b1) {
  return 2.0 * a0 * b0 + (2.0 / 3.0) * a1 * b1;
}

inline double int_terms(const Term& a, const Term& b) {
  double v = a.coef * b.coef;
  for (int k = 0; k < 3; ++k) v *= int1(a.p0[k], a.p1[k], b.p0[k], b.p1[k]);
  return v;
}

inline int levi(int a, int b, int c) {
  if (a == b || b == c || a == c) return 0;
  // even permutations of (0,1,2): (0,1,2), (1,2,0), (2,0,1)
  if ((a == 0 && b == 1) || (a == 1 && b == 2) || (a == 2 && b == 0)) return 1;
  return -1;
}

// Element matrices of the scaled lowest-order edge basis on a brick hx*hy*hz.
// basis_i = (2/h_d) * 1/4 * prod_{k!=d} (1 + s_k xi_k) * e_d ; b = int curl.curl ; c = int basis.basis
void element_matrices(double hx, double hy, double hz, double b[12][12], double c[12][12]) {
  const double h[3] = {hx, hy, hz};
  const double detj = hx * hy * hz / 8.0;
  // mass term: a single component (direction d)
  Term phi[12];
  // curl terms: component m, one term each (zero when m == d)
  Term curl[12][3];
  bool curl_nz[12][3];
  for (int i = 0; i < 12; ++i) {
    const int d = kLocalDir[i];
    const double base = 0.5 / h[d];  // (2/h_d) * 1/4
    phi[i].coef = base;
    for (int k = 0; k < 3; ++k) {
      phi[i].p0[k] = 1.0;
      phi[i].p1[k] = (k == d) ? 0.0 : double(kLocalSign[i][k]);
    }
    for (int m = 0; m < 3; ++m) {
      curl_nz[i][m] = false;
      if (m == d) continue;
      const int l = 3 - d - m;  // the remaining coordinate
      const int eps = levi(l, d, m);
      Term t;
      t.coef = eps * base * (2.0 / h[l]) * double(kLocalSign[i][l]);
      for (int k = 0; k < 3; ++k) {
        t.p0[k] = 1.0;
        t.p1[k] = (k == m) ? double(kLocalSign[i][m]) : 0.0;
      }
      curl[i][m] = t;
      curl_nz[i][m] = true;
    }
  }
  for (int i = 0; i < 12; ++i) {
    for (int j = 0; j < 12; ++j) {
      double bij = 0.0;
      for (int m = 0; m < 3; ++m)
        if (curl_nz[i][m] && curl_nz[j][m]) bij += int_terms(curl[i][m], curl[j][m]);
      b[i][j] = bij * detj;
      c[i][j] = (kLocalDir[i] == kLocalDir[j]) ? int_terms(phi[i], phi[j]) * detj : 0.0;
    }
  }
}

// Linear interpolation with end clamping (Code/BuildMatrix/For_Solvers.cpp:227-266): the first
// interval [xs[i], xs[i+1]] that contains x, found by bisection instead of a linear scan.
double spline(double x, long long n, const double* xs, const double* v) {
  if (x < xs[0]) return v[0];
  if (x > xs[n - 1]) return v[n - 1];
  // invariant: xs[lo] <= x <= xs[hi]; lo ends as the last index with xs[lo] < x (or 0), which is
  // the first bracketing interval of the scan
  long long lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const long long mid = (lo + hi) / 2;
    if (xs[mid] < x) lo = mid; else hi = mid;
  }
  const double t = (x - xs[lo]) / (xs[lo + 1] - xs[lo]);
  return (1.0 - t) * v[lo] + t * v[lo + 1];
}

// Cells that contain a given edge, ascending cell order, and the edge's local index there.
int edge_cells(const Grid& g, int dir, long long ix, long long iy, long long iz, long long* cells,
               int* local) {
  int n = 0;
  const long long ncx = g.nxn - 1, ncy = g.nyn - 1, ncz = g.nzn - 1;
  auto cell_id = [&](long long cx, long long cy, long long cz) {
    return (cz * ncy + cy) * ncx + cx;
  };
  if (dir == 0) {  // x-edge at (y = iy, z = iz): cells (ix, iy-1|iy, iz-1|iz)
    for (int dz = -1; dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy) {
        const long long cy = iy + dy, cz = iz + dz;
        if (cy < 0 || cy >= ncy || cz < 0 || cz >= ncz) continue;
        cells[n] = cell_id(ix, cy, cz);
        local[n] = (dy == 0 ? 0 : 1) + (dz == 0 ? 0 : 2);  // position inside the cell
        ++n;
      }
  } else if (dir == 1) {  // y-edge at (x = ix, z = iz): local 4 (x-,z-) 5 (x-,z+) 6 (x+,z-) 7 (x+,z+)
    for (int dz = -1; dz <= 0; ++dz)
      for (int dx = -1; dx <= 0; ++dx) {
        const long long cx = ix + dx, cz = iz + dz;
        if (cx < 0 || cx >= ncx || cz < 0 || cz >= ncz) continue;
        cells[n] = cell_id(cx, iy, cz);
        local[n] = 4 + (dx == 0 ? 0 : 2) + (dz == 0 ? 0 : 1);
        ++n;
      }
  } else {  // z-edge at (x = ix, y = iy): local 8 (x-,y-) 9 (x+,y-) 10 (x-,y+) 11 (x+,y+)
    for (int dy = -1; dy <= 0; ++dy)
      for (int dx = -1; dx <= 0; ++dx) {
        const long long cx = ix + dx, cy = iy + dy;
        if (cx < 0 || cx >= ncx || cy < 0 || cy >= ncy) continue;
        cells[n] = cell_id(cx, cy, iz);
        local[n] = 8 + (dx == 0 ? 0 : 1) + (dy == 0 ? 0 : 2);
        ++n;
      }
  }
  return n;  // cells are produced in ascending order
}

// Geometric description of every edge: first node (ix,iy,iz) and direction.
struct EdgeGeom {
  std::vector<int> ix, iy, iz;
  std::vector<signed char> dir;
};

void build_edge_geom(const Grid& g, EdgeGeom& eg) {
  const long long nb = g.edges();
  eg.ix.resize(nb);
  eg.iy.resize(nb);
  eg.iz.resize(nb);
  eg.dir.resize(nb);
  long long e = 0;
  for (long long iz = 0; iz < g.nzn; ++iz)
    for (long long iy = 0; iy < g.nyn; ++iy)
      for (long long ix = 0; ix < g.nxn; ++ix) {
        if (ix < g.nxn - 1) { eg.ix[e] = ix; eg.iy[e] = iy; eg.iz[e] = iz; eg.dir[e] = 0; ++e; }
        if (iy < g.nyn - 1) { eg.ix[e] = ix; eg.iy[e] = iy; eg.iz[e] = iz; eg.dir[e] = 1; ++e; }
        if (iz < g.nzn - 1) { eg.ix[e] = ix; eg.iy[e] = iy; eg.iz[e] = iz; eg.dir[e] = 2; ++e; }
      }
}

// Lower neighbours (edges sharing a cell, index < e) of edge e, ascending and unique.
int lower_neighbours(const Grid& g, const EdgeNumbering& num, const EdgeGeom& eg, long long e,
                     long long* out) {
  long long cells[4];
  int local[4];
  const int nc = edge_cells(g, eg.dir[e], eg.ix[e], eg.iy[e], eg.iz[e], cells, local);
  const long long ncx = g.nxn - 1, ncy = g.nyn - 1;
  long long cand[48];
  int n = 0;
  for (int c = 0; c < nc; ++c) {
    const long long cid = cells[c];
    const long long cx = cid % ncx, cy = (cid / ncx) % ncy, cz = cid / (ncx * ncy);
    long long ce[12];
    num.cell_edges(cx, cy, cz, ce);
    for (int k = 0; k < 12; ++k)
      if (ce[k] < e) cand[n++] = ce[k];
  }
  std::sort(cand, cand + n);
  const int m = int(std::unique(cand, cand + n) - cand);
  for (int k = 0; k < m; ++k) out[k] = cand[k];
  return m;
}

bool is_boundary_node(const Grid& g, long long ix, long long iy, long long iz) {
  return ix == 0 || iy == 0 || iz == 0 || ix == g.nxn - 1 || iy == g.nyn - 1 || iz == g.nzn - 1;
}

// Boundary edge = both end nodes on the outer boundary (bound_cond_vec_harm.cpp:81-108).
bool is_boundary_edge(const Grid& g, const EdgeGeom& eg, long long e) {
  const long long ix = eg.ix[e], iy = eg.iy[e], iz = eg.iz[e];
  const int d = eg.dir[e];
  const long long jx = ix + (d == 0), jy = iy + (d == 1), jz = iz + (d == 2);
  return is_boundary_node(g, ix, iy, iz) && is_boundary_node(g, jx, jy, jz);
}

// Position of column jj inside the lower row ii (ascending jg), or -1.
inline long long find_in_row(const int* ig, const int* jg, long long ii, long long jj) {
  const int* b = jg + ig[ii];
  const int* e = jg + ig[ii + 1];
  const int* p = std::lower_bound(b, e, int(jj));
  if (p != e && *p == jj) return ig[ii] + (p - b);
  return -1;
}

// Round through the reference's "%25.13e" text exchange (Harm1D writes, BuildMatrix reads).
double round_text(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%25.13e", v);
  return std::strtod(buf, nullptr);
}

}  // namespace

extern "C" {

int64_t mtmodel_num_edges(int nxn, int nyn, int nzn) {
  if (!grid_ok(nxn, nyn, nzn)) return MTMODEL_ERR_ARG;
  Grid g{nxn, nyn, nzn};
  return g.edges();
}

int mtmodel_topology(int nxn, int nyn, int nzn, int* nver, int* ed, int* nodes_for_edges,
                     int* l13d, int64_t* kt1) {
  if (!grid_ok(nxn, nyn, nzn)) return MTMODEL_ERR_ARG;
  Grid g{nxn, nyn, nzn};
  EdgeNumbering num(g);
  const long long ncx = g.nxn - 1, ncy = g.nyn - 1, ncz = g.nzn - 1;
  if (g.edges() > 2147483000LL || g.nodes() > 2147483000LL) return MTMODEL_ERR_SIZE;
  if (nver || ed) {
    long long cell = 0;
    for (long long cz = 0; cz < ncz; ++cz)
      for (long long cy = 0; cy < ncy; ++cy)
        for (long long cx = 0; cx < ncx; ++cx, ++cell) {
          if (nver) {
            int* v = nver + cell * 14;
            const long long n = g.node(cx, cy, cz) + 1;  // 1-based
            v[0] = int(n);
            v[1] = int(n + 1);
            v[2] = int(n + g.nxn);
            v[3] = int(n + g.nxn + 1);
            v[4] = int(n + g.nxy());
            v[5] = int(n + g.nxy() + 1);
            v[6] = int(n + g.nxy() + g.nxn);
            v[7] = int(n + g.nxy() + g.nxn + 1);
            for (int k = 8; k < 14; ++k) v[k] = 0;
          }
          if (ed) {
            long long ce[12];
            num.cell_edges(cx, cy, cz, ce);
            int* r = ed + cell * 25;
            for (int k = 0; k < 12; ++k) r[k] = int(ce[k] + 1);
            for (int k = 12; k < 24; ++k) r[k] = 0;
            r[24] = 1;
          }
        }
  }
  if (nodes_for_edges) {
    long long e = 0;
    for (long long iz = 0; iz < g.nzn; ++iz)
      for (long long iy = 0; iy < g.nyn; ++iy)
        for (long long ix = 0; ix < g.nxn; ++ix) {
          const long long n = g.node(ix, iy, iz) + 1;
          if (ix < g.nxn - 1) { nodes_for_edges[2 * e] = int(n); nodes_for_edges[2 * e + 1] = int(n + 1); ++e; }
          if (iy < g.nyn - 1) { nodes_for_edges[2 * e] = int(n); nodes_for_edges[2 * e + 1] = int(n + g.nxn); ++e; }
          if (iz < g.nzn - 1) { nodes_for_edges[2 * e] = int(n); nodes_for_edges[2 * e + 1] = int(n + g.nxy()); ++e; }
        }
  }
  long long k = 0;
  for (long long iz = 0; iz < g.nzn; ++iz)
    for (long long iy = 0; iy < g.nyn; ++iy)
      for (long long ix = 0; ix < g.nxn; ++ix)
        if (is_boundary_node(g, ix, iy, iz)) {
          if (l13d) l13d[k] = int(g.node(ix, iy, iz) + 1);
          ++k;
        }
  if (kt1) *kt1 = k;
  return 0;
}

int mtmodel_portrait_size(int nxn, int nyn, int nzn, int64_t* nnz_lower) {
  if (!grid_ok(nxn, nyn, nzn) || !nnz_lower) return MTMODEL_ERR_ARG;
  Grid g{nxn, nyn, nzn};
  EdgeNumbering num(g);
  EdgeGeom eg;
  build_edge_geom(g, eg);
  const long long nb = g.edges();
  long long total = 0;
#pragma omp parallel for reduction(+ : total) schedule(static)
  for (long long e = 0; e < nb; ++e) {
    long long buf[48];
    total += lower_neighbours(g, num, eg, e, buf);
  }
  *nnz_lower = total;
  if (total > 2147483000LL) return MTMODEL_ERR_SIZE;
  return 0;
}

int mtmodel_portrait(int nxn, int nyn, int nzn, int* ig, int* jg, int* idi, int* ijg) {
  if (!grid_ok(nxn, nyn, nzn) || !ig || !jg || !idi || !ijg) return MTMODEL_ERR_ARG;
  Grid g{nxn, nyn, nzn};
  EdgeNumbering num(g);
  EdgeGeom eg;
  build_edge_geom(g, eg);
  const long long nb = g.edges();
  std::vector<int> cnt(nb);
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nb; ++e) {
    long long buf[48];
    cnt[e] = lower_neighbours(g, num, eg, e, buf);
  }
  ig[0] = 0;
  for (long long e = 0; e < nb; ++e) {
    const long long next = (long long)ig[e] + cnt[e];
    if (next > 2147483000LL) return MTMODEL_ERR_SIZE;
    ig[e + 1] = int(next);
  }
  // block sizes: 2 for parallel edge pairs (and all diagonals), 1 for orthogonal pairs
  std::vector<signed char> bsz(ig[nb]);
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nb; ++e) {
    long long buf[48];
    const int n = lower_neighbours(g, num, eg, e, buf);
    for (int k = 0; k < n; ++k) {
      jg[ig[e] + k] = int(buf[k]);
      bsz[ig[e] + k] = (eg.dir[buf[k]] == eg.dir[e]) ? 2 : 1;
    }
  }
  for (long long e = 0; e <= nb; ++e) idi[e] = int(2 * e);
  ijg[0] = 0;
  for (long long k = 0; k < ig[nb]; ++k) {
    const long long next = (long long)ijg[k] + bsz[k];
    if (next > 2147483000LL) return MTMODEL_ERR_SIZE;
    ijg[k + 1] = int(next);
  }
  return 0;
}

int mtmodel_element_matrices(double hx, double hy, double hz, double* b, double* c) {
  if (!(hx > 0 && hy > 0 && hz > 0) || !b || !c) return MTMODEL_ERR_ARG;
  double bb[12][12], cc[12][12];
  element_matrices(hx, hy, hz, bb, cc);
  std::memcpy(b, bb, sizeof(bb));
  std::memcpy(c, cc, sizeof(cc));
  return 0;
}

int mtmodel_harm1d(int n_layers, const double* layer_top, const double* layer_sigma, double nu,
                   int max_points, int* n_points, double* z1d, double* usin, double* ucos) {
  if (n_layers < 1 || !layer_top || !layer_sigma || !n_points || !(nu > 0)) return MTMODEL_ERR_ARG;
  const double x0_1d = -10000000.0, x1_1d = 0.0, coef = 1.02, step0 = 1e-2;
  std::vector<double> layers(n_layers + 1);
  layers[0] = x0_1d;
  for (int i = 0; i < n_layers; ++i) layers[i + 1] = layer_top[i];
  // node placement (Mesh_1d_mtz::Gen_1d_mesh): the step keeps growing across layer interfaces
  // and the running coordinate is not reset to the interface after it is crossed.
  std::vector<double> coords;
  {
    double x = x1_1d, step = step0;
    for (int i = 0; i < n_layers; ++i) {
      coords.push_back(layers[n_layers - i]);
      while (true) {
        x -= step;
        if (x <= layers[n_layers - i - 1]) break;
        step *= coef;
        coords.push_back(x);
      }
    }
    coords.push_back(x0_1d);
    std::reverse(coords.begin(), coords.end());
  }
  const long long np = (long long)coords.size();
  *n_points = int(np);
  if (np > max_points || !z1d || !usin || !ucos) return np > max_points ? MTMODEL_ERR_SIZE : 0;
  const long long ne = np - 1;
  const double omega = nu * (2.0 * kPi);
  const double mu = 4.0 * kPi * 1E-7;
  // complex tridiagonal system (b + i c) per element, b = 1/(mu h) stiffness, c = sigma omega h/6
  std::vector<double> dr(np, 0.0), di(np, 0.0), lr(np, 0.0), li(np, 0.0), ur(np, 0.0), ui(np, 0.0);
  for (long long e = 0; e < ne; ++e) {
    const double h = coords[e + 1] - coords[e];
    const double center = 0.5 * (coords[e] + coords[e + 1]);
    double sig = layer_sigma[0];
    for (int j = 0; j < n_layers; ++j)
      if (center > layers[j] && center < layers[j + 1]) { sig = layer_sigma[j]; break; }
    const double tb = 1.0 / (mu * h);
    const double tc = sig * omega * h / 6.0;
    dr[e] += tb; di[e] += 2.0 * tc;
    dr[e + 1] += tb; di[e + 1] += 2.0 * tc;
    ur[e] += -tb; ui[e] += tc;          // (e, e+1)
    lr[e + 1] += -tb; li[e + 1] += tc;  // (e+1, e)
  }
  // Dirichlet A = 0 at z = -1e7; unit natural flux at the surface node
  std::vector<double> rr(np, 0.0), ri(np, 0.0);
  dr[0] = 1.0; di[0] = 0.0; ur[0] = 0.0; ui[0] = 0.0;
  lr[1] = 0.0; li[1] = 0.0;
  rr[np - 1] += 1.0;
  // Thomas algorithm in complex arithmetic (no pivoting, like the reference's profile LU)
  std::vector<double> cpr(np), cpi(np), dpr(np), dpi(np);
  auto cdiv = [](double ar, double ai, double br, double bi, double& qr, double& qi) {
    const double den = br * br + bi * bi;
    qr = (ar * br + ai * bi) / den;
    qi = (ai * br - ar * bi) / den;
  };
  {
    cdiv(ur[0], ui[0], dr[0], di[0], cpr[0], cpi[0]);
    cdiv(rr[0], ri[0], dr[0], di[0], dpr[0], dpi[0]);
    for (long long k = 1; k < np; ++k) {
      // m = d_k - l_k * cp_{k-1}
      const double mr = dr[k] - (lr[k] * cpr[k - 1] - li[k] * cpi[k - 1]);
      const double mi = di[k] - (lr[k] * cpi[k - 1] + li[k] * cpr[k - 1]);
      if (mr == 0.0 && mi == 0.0) return MTMODEL_ERR_SOLVE;
      if (k < np - 1) cdiv(ur[k], ui[k], mr, mi, cpr[k], cpi[k]);
      const double nr = rr[k] - (lr[k] * dpr[k - 1] - li[k] * dpi[k - 1]);
      const double ni = ri[k] - (lr[k] * dpi[k - 1] + li[k] * dpr[k - 1]);
      cdiv(nr, ni, mr, mi, dpr[k], dpi[k]);
    }
    std::vector<double> xr(np), xi(np);
    xr[np - 1] = dpr[np - 1];
    xi[np - 1] = dpi[np - 1];
    for (long long k = np - 2; k >= 0; --k) {
      xr[k] = dpr[k] - (cpr[k] * xr[k + 1] - cpi[k] * xi[k + 1]);
      xi[k] = dpi[k] - (cpr[k] * xi[k + 1] + cpi[k] * xr[k + 1]);
    }
    for (long long k = 0; k < np; ++k) {
      z1d[k] = round_text(coords[k]);
      usin[k] = round_text(xr[k]);
      ucos[k] = round_text(xi[k]);
    }
  }
  return 0;
}

}  // extern "C"

namespace {

// Shared assembly loop: matrix (when di/gg are given) and right-hand side.
int assemble_impl(int nxn, int nyn, int nzn, const double* x, const double* y, const double* z,
                  const double* cell_sigma, const double* cell_sigma0, const double* cell_mu,
                  double nu, int alpha, int n1d, const double* z1d, const double* usin,
                  const double* ucos, const int* ig, const int* jg, const int* idi,
                  const int* ijg, double* di, double* gg, double* pr) {
  const bool want_matrix = (di != nullptr);
  Grid g{nxn, nyn, nzn};
  EdgeNumbering num(g);
  EdgeGeom eg;
  build_edge_geom(g, eg);
  const long long nb = g.edges();
  const long long ncx = g.nxn - 1, ncy = g.nyn - 1;
  const double omega = 2.0 * kPi * nu;
  const double alpha_d = double(alpha);
  // Row-parallel gather: each matrix entry and RHS entry receives its element contributions in
  // ascending cell order, which is the summation order of the reference's element loop.
#pragma omp parallel for schedule(dynamic, 4096)
  for (long long e = 0; e < nb; ++e) {
    long long cells[4];
    int local[4];
    const int nc = edge_cells(g, eg.dir[e], eg.ix[e], eg.iy[e], eg.iz[e], cells, local);
    double d0 = 0.0, d1 = 0.0, p0 = 0.0, p1 = 0.0;
    if (want_matrix)
      for (long long m = ig[e]; m < ig[e + 1]; ++m)
        for (int k = ijg[m]; k < ijg[m + 1]; ++k) gg[k] = 0.0;
    for (int c = 0; c < nc; ++c) {
      const long long cid = cells[c];
      const long long cx = cid % ncx, cy = (cid / ncx) % ncy, cz = cid / (ncx * ncy);
      const double hx = x[cx + 1] - x[cx], hy = y[cy + 1] - y[cy], hz = z[cz + 1] - z[cz];
      double b[12][12], cm[12][12];
      element_matrices(hx, hy, hz, b, cm);
      long long ce[12];
      num.cell_edges(cx, cy, cz, ce);
      const double sigma = cell_sigma[cid], sigma0 = cell_sigma0[cid], mu = cell_mu[cid];
      const int j = local[c];
      // RHS (T_Brick::Calc_block_local_matrix_and_vector / Calc_local_vector_for_MT) with the
      // unscaled mass matrix and the normal field's edge DOFs at edge midpoints.
      double asin0[12], acos0[12];
      const double hloc[3] = {hx, hy, hz};
      const double zc0 = z[cz], zc1 = z[cz + 1];
      for (int i = 0; i < 12; ++i) {
        const int d = kLocalDir[i];
        double zm;
        if (d == 2) zm = 0.5 * (zc0 + zc1);
        else zm = (kLocalSign[i][2] < 0) ? 0.5 * (zc0 + zc0) : 0.5 * (zc1 + zc1);
        const double vs = spline(zm, n1d, z1d, usin);
        const double vc = spline(zm, n1d, z1d, ucos);
        // u = (v*alpha, v*(1-alpha), 0), DOF = u . (J t) with J = diag(h/2)
        double us[3] = {vs * alpha_d, vs * (1.0 - alpha_d), 0.0};
        double uc[3] = {vc * alpha_d, vc * (1.0 - alpha_d), 0.0};
        double jt[3] = {0.0, 0.0, 0.0};
        jt[d] = hloc[d] * 0.5;
        asin0[i] = us[0] * jt[0] + us[1] * jt[1] + us[2] * jt[2];
        acos0[i] = uc[0] * jt[0] + uc[1] * jt[1] + uc[2] * jt[2];
      }
      double g_re = 0.0, g_im = 0.0;
      for (int k = 0; k < 12; ++k) g_re += cm[j][k] * asin0[k];
      for (int k = 0; k < 12; ++k) g_im += cm[j][k] * acos0[k];
      const double mult = (sigma0 - sigma) * omega;
      p0 += g_im * (-mult);
      p1 += g_re * mult;
      if (!want_matrix) continue;
      // scaled element matrix row j: b/mu (real), c*sigma*omega (imaginary)
      const double so = sigma * omega;
      d0 += b[j][j] / mu;
      d1 += cm[j][j] * so;
      for (int k = 0; k < 12; ++k) {
        const long long jj = ce[k];
        if (jj >= e) continue;
        const long long pos = find_in_row(ig, jg, e, jj);
        if (pos < 0) continue;
        const int beg = ijg[pos];
        gg[beg] += b[j][k] / mu;
        if (ijg[pos + 1] - beg == 2) gg[beg + 1] += cm[j][k] * so;
      }
    }
    if (want_matrix) {
      di[idi[e]] = d0;
      if (idi[e + 1] - idi[e] == 2) di[idi[e] + 1] = d1;
    }
    pr[2 * e] = p0;
    pr[2 * e + 1] = p1;
  }
  // homogeneous Dirichlet on boundary edges, then symmetrisation of the columns
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nb; ++e) {
    if (is_boundary_edge(g, eg, e)) {
      pr[2 * e] = pr[2 * e + 1] = 0.0;
      if (!want_matrix) continue;
      di[idi[e]] = 1.0;
      if (idi[e + 1] - idi[e] == 2) di[idi[e] + 1] = 0.0;
      for (long long m = ig[e]; m < ig[e + 1]; ++m)
        for (int k = ijg[m]; k < ijg[m + 1]; ++k) gg[k] = 0.0;
    }
    if (!want_matrix) continue;
    for (long long m = ig[e]; m < ig[e + 1]; ++m)
      if (is_boundary_edge(g, eg, jg[m]))
        for (int k = ijg[m]; k < ijg[m + 1]; ++k) gg[k] = 0.0;
  }
  return 0;
}

}  // namespace

extern "C" {

int mtmodel_assemble(int nxn, int nyn, int nzn, const double* x, const double* y, const double* z,
                     const double* cell_sigma, const double* cell_sigma0, const double* cell_mu,
                     double nu, int alpha, int n1d, const double* z1d, const double* usin,
                     const double* ucos, const int* ig, const int* jg, const int* idi,
                     const int* ijg, double* di, double* gg, double* pr) {
  if (!grid_ok(nxn, nyn, nzn) || !x || !y || !z || !cell_sigma || !cell_sigma0 || !cell_mu ||
      !ig || !jg || !idi || !ijg || !di || !gg || !pr || n1d < 2 || !z1d || !usin || !ucos)
    return MTMODEL_ERR_ARG;
  return assemble_impl(nxn, nyn, nzn, x, y, z, cell_sigma, cell_sigma0, cell_mu, nu, alpha, n1d,
                       z1d, usin, ucos, ig, jg, idi, ijg, di, gg, pr);
}

int mtmodel_assemble_rhs(int nxn, int nyn, int nzn, const double* x, const double* y,
                         const double* z, const double* cell_sigma, const double* cell_sigma0,
                         double nu, int alpha, int n1d, const double* z1d, const double* usin,
                         const double* ucos, double* pr) {
  if (!grid_ok(nxn, nyn, nzn) || !x || !y || !z || !cell_sigma || !cell_sigma0 || !pr ||
      n1d < 2 || !z1d || !usin || !ucos)
    return MTMODEL_ERR_ARG;
  // Same arithmetic as the assembly loop, with each cell's 12-entry local vector computed once
  // and then gathered per edge in ascending cell order.
  Grid g{nxn, nyn, nzn};
  EdgeGeom eg;
  build_edge_geom(g, eg);
  const long long nb = g.edges();
  const long long ncx = g.nxn - 1, ncy = g.nyn - 1, ncz = g.nzn - 1;
  const long long ncell = ncx * ncy * ncz;
  const double omega = 2.0 * kPi * nu;
  const double alpha_d = double(alpha);
  std::vector<double> loc((size_t)ncell * 24);
#pragma omp parallel for schedule(static)
  for (long long cid = 0; cid < ncell; ++cid) {
    const long long cx = cid % ncx, cy = (cid / ncx) % ncy, cz = cid / (ncx * ncy);
    const double hx = x[cx + 1] - x[cx], hy = y[cy + 1] - y[cy], hz = z[cz + 1] - z[cz];
    double* out = &loc[(size_t)cid * 24];
    const double sigma = cell_sigma[cid], sigma0 = cell_sigma0[cid];
    if (sigma == sigma0) {  // mult = 0: every entry is +-0
      for (int k = 0; k < 24; ++k) out[k] = 0.0;
      continue;
    }
    double b[12][12], cm[12][12];
    element_matrices(hx, hy, hz, b, cm);
    double asin0[12], acos0[12];
    const double hloc[3] = {hx, hy, hz};
    const double zc0 = z[cz], zc1 = z[cz + 1];
    for (int i = 0; i < 12; ++i) {
      const int d = kLocalDir[i];
      double zm;
      if (d == 2) zm = 0.5 * (zc0 + zc1);
      else zm = (kLocalSign[i][2] < 0) ? 0.5 * (zc0 + zc0) : 0.5 * (zc1 + zc1);
      const double vs = spline(zm, n1d, z1d, usin);
      const double vc = spline(zm, n1d, z1d, ucos);
      double us[3] = {vs * alpha_d, vs * (1.0 - alpha_d), 0.0};
      double uc[3] = {vc * alpha_d, vc * (1.0 - alpha_d), 0.0};
      double jt[3] = {0.0, 0.0, 0.0};
      jt[d] = hloc[d] * 0.5;
      asin0[i] = us[0] * jt[0] + us[1] * jt[1] + us[2] * jt[2];
      acos0[i] = uc[0] * jt[0] + uc[1] * jt[1] + uc[2] * jt[2];
    }
    const double mult = (sigma0 - sigma) * omega;
    for (int j = 0; j < 12; ++j) {
      double g_re = 0.0, g_im = 0.0;
      for (int k = 0; k < 12; ++k) g_re += cm[j][k] * asin0[k];
      for (int k = 0; k < 12; ++k) g_im += cm[j][k] * acos0[k];
      out[2 * j] = g_im * (-mult);
      out[2 * j + 1] = g_re * mult;
    }
  }
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nb; ++e) {
    if (is_boundary_edge(g, eg, e)) {
      pr[2 * e] = pr[2 * e + 1] = 0.0;
      continue;
    }
    long long cells[4];
    int local[4];
    const int nc = edge_cells(g, eg.dir[e], eg.ix[e], eg.iy[e], eg.iz[e], cells, local);
    double p0 = 0.0, p1 = 0.0;
    for (int c = 0; c < nc; ++c) {
      p0 += loc[(size_t)cells[c] * 24 + 2 * local[c]];
      p1 += loc[(size_t)cells[c] * 24 + 2 * local[c] + 1];
    }
    pr[2 * e] = p0;
    pr[2 * e + 1] = p1;
  }
  return 0;
}

int mtmodel_assemble_ks(int nxn, int nyn, int nzn, const double* x, const double* y,
                        const double* z, const double* cell_sigma, const double* cell_mu,
                        const int* ig, const int* jg, double* k_di, double* k_gg, double* s_di,
                        double* s_gg) {
  if (!grid_ok(nxn, nyn, nzn) || !x || !y || !z || !cell_sigma || !cell_mu || !ig || !jg ||
      !k_di || !k_gg || !s_di || !s_gg)
    return MTMODEL_ERR_ARG;
  Grid g{nxn, nyn, nzn};
  EdgeNumbering num(g);
  EdgeGeom eg;
  build_edge_geom(g, eg);
  const long long nb = g.edges();
  const long long ncx = g.nxn - 1, ncy = g.nyn - 1;
#pragma omp parallel for schedule(dynamic, 4096)
  for (long long e = 0; e < nb; ++e) {
    long long cells[4];
    int local[4];
    const int nc = edge_cells(g, eg.dir[e], eg.ix[e], eg.iy[e], eg.iz[e], cells, local);
    for (long long m = ig[e]; m < ig[e + 1]; ++m) k_gg[m] = s_gg[m] = 0.0;
    double kd = 0.0, sd = 0.0;
    for (int c = 0; c < nc; ++c) {
      const long long cid = cells[c];
      const long long cx = cid % ncx, cy = (cid / ncx) % ncy, cz = cid / (ncx * ncy);
      double b[12][12], cm[12][12];
      element_matrices(x[cx + 1] - x[cx], y[cy + 1] - y[cy], z[cz + 1] - z[cz], b, cm);
      long long ce[12];
      num.cell_edges(cx, cy, cz, ce);
      const int j = local[c];
      const double mu = cell_mu[cid], sigma = cell_sigma[cid];
      kd += b[j][j] / mu;
      sd += cm[j][j] * sigma;
      for (int k = 0; k < 12; ++k) {
        const long long jj = ce[k];
        if (jj >= e) continue;
        const long long pos = find_in_row(ig, jg, e, jj);
        if (pos < 0) continue;
        k_gg[pos] += b[j][k] / mu;
        s_gg[pos] += cm[j][k] * sigma;
      }
    }
    k_di[e] = kd;
    s_di[e] = sd;
  }
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nb; ++e) {
    const bool be = is_boundary_edge(g, eg, e);
    if (be) {
      k_di[e] = 1.0;
      s_di[e] = 0.0;
    }
    for (long long m = ig[e]; m < ig[e + 1]; ++m)
      if (be || is_boundary_edge(g, eg, jg[m])) k_gg[m] = s_gg[m] = 0.0;
  }
  return 0;
}

int mtmodel_normal_dofs(int nxn, int nyn, int nzn, const double* x, const double* y,
                        const double* z, int alpha, int n1d, const double* z1d,
                        const double* usin, const double* ucos, double* a0) {
  if (!grid_ok(nxn, nyn, nzn) || !x || !y || !z || !z1d || !usin || !ucos || !a0 || n1d < 2)
    return MTMODEL_ERR_ARG;
  Grid g{nxn, nyn, nzn};
  EdgeGeom eg;
  build_edge_geom(g, eg);
  const long long nb = g.edges();
  const double alpha_d = double(alpha);
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nb; ++e) {
    const int d = eg.dir[e];
    double re = 0.0, im = 0.0;
    if (d != 2) {
      // the DOF the right-hand side uses: (u . t) h / 2 at the edge's height, u = A0 (alpha,
      // 1 - alpha, 0) (see assemble_impl / mtmodel_assemble_rhs)
      const double zm = 0.5 * (z[eg.iz[e]] + z[eg.iz[e]]);
      const double h = d == 0 ? x[eg.ix[e] + 1] - x[eg.ix[e]] : y[eg.iy[e] + 1] - y[eg.iy[e]];
      const double w = d == 0 ? alpha_d : 1.0 - alpha_d;
      re = spline(zm, n1d, z1d, usin) * w * (h * 0.5);
      im = spline(zm, n1d, z1d, ucos) * w * (h * 0.5);
    }
    a0[2 * e] = re;
    a0[2 * e + 1] = im;
  }
  return 0;
}

int mtmodel_cell_mass_forms(int nxn, int nyn, int nzn, const double* x, const double* y,
                            const double* z, const double* u, const double* v, double* out) {
  if (!grid_ok(nxn, nyn, nzn) || !x || !y || !z || !u || !v || !out) return MTMODEL_ERR_ARG;
  Grid g{nxn, nyn, nzn};
  EdgeNumbering num(g);
  const long long ncx = g.nxn - 1, ncy = g.nyn - 1, ncz = g.nzn - 1;
#pragma omp parallel for schedule(static)
  for (long long cid = 0; cid < ncx * ncy * ncz; ++cid) {
    const long long cx = cid % ncx, cy = (cid / ncx) % ncy, cz = cid / (ncx * ncy);
    double b[12][12], cm[12][12];
    element_matrices(x[cx + 1] - x[cx], y[cy + 1] - y[cy], z[cz + 1] - z[cz], b, cm);
    long long ce[12];
    num.cell_edges(cx, cy, cz, ce);
    double sr = 0.0, si = 0.0;  // sum_jk u_j M_jk v_k, complex, unconjugated
    for (int j = 0; j < 12; ++j) {
      const double ur = u[2 * ce[j]], ui = u[2 * ce[j] + 1];
      double tr = 0.0, ti = 0.0;
      for (int k = 0; k < 12; ++k) {
        tr += cm[j][k] * v[2 * ce[k]];
        ti += cm[j][k] * v[2 * ce[k] + 1];
      }
      sr += ur * tr - ui * ti;
      si += ur * ti + ui * tr;
    }
    out[2 * cid] = sr;
    out[2 * cid + 1] = si;
  }
  return 0;
}

}  // extern "C"
