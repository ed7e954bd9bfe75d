// Host-side preparation of the solver's device layout.
//
// SlicedEll: the full complex-symmetric matrix (both triangles, explicit zero blocks dropped)
// in 64-row slices. Each slice stores its rows' entries column-major (slot k of row r at
// off + k*64 + (r % 64)) so a wavefront's 64 lanes read 64 consecutive entries. Entries are
// split into a "complex" section (col, a, b) meaning a + i*omega*b, and a "real" section
// (col, a). A single system from the reference's files uses omega = 1 (a = Re, b = Im); a
// frequency-independent pair uses a = K, b = S and A(omega) = K + i*omega*S.
//
// GradientEll: the folded preconditioner's discrete gradient G (interior nodes x edges) and its
// transpose, padded row-wise (column-major) to their maximum row length.
#pragma once

#include <cstdint>
#include <vector>

namespace mtcocr {

struct SlicedEll {
  int nb = 0;
  int nslices = 0;
  std::vector<int> offc, nc, offr, nr;  // per slice: entry offset and slot count
  std::vector<int> colc, colr;
  std::vector<double> ac, bc, ar;
  std::vector<double> diag_a, diag_b;   // diagonal as a + i*omega*b
  int64_t nnz_complex = 0, nnz_real = 0;  // stored (unpadded) entries
  // Row pack: the off-diagonal entries row-contiguous (no padding) for the row-group SpMV,
  // where the lanes of one row load consecutive entries together; the diagonal is diag_a/b.
  // Row i's entries are [rp[i], rp[i+1]) of pcol/pa: first its complex entries (ascending
  // columns), whose b values are [rb[i], rb[i+1]) of pb, then its real ones (ascending).
  std::vector<int> rp, rb, pcol;
  std::vector<double> pa, pb;
  // Row blocks for the halo-staged SpMV: block k is rows [16k, 16k + 16); its distinct columns
  // (the rows themselves included) are halo[hoff[k] .. hoff[k+1]), ascending. Every row-pack
  // entry also has its column's position in its block's halo (lcol, parallel to pcol), every
  // row its own position (ldiag). halo_max is the largest halo; blocks are not built (hoff
  // empty) when one would exceed kHaloCap columns.
  static constexpr int kBlockRows = 16;
  static constexpr int kHaloCap = 512;
  std::vector<int> hoff, halo;
  std::vector<unsigned short> lcol, ldiag;
  int halo_max = 0;
};

struct GradientEll {
  int n = 0;   // edges (rows of G^T)
  int m = 0;   // interior nodes (rows of G)
  int wg = 0;  // max entries per G row
  int wgt = 0; // max entries per G^T row
  std::vector<int> gcol;     // [k * m + i]
  std::vector<double> gval;
  std::vector<int> gtcol;    // [k * n + e]
  std::vector<double> gtval;
};

// From the reference's lower-triangle 2x2-block arrays (0-based). `order` (new -> old row,
// may be null) permutes rows and columns symmetrically; every row keeps its entry order.
// Returns 0 or a negative code.
int ell_from_block2x2(int nb, const int* ig, const int* jg, const int* idi, const int* ijg,
                      const double* di, const double* gg, const int* order, SlicedEll& out);

// From a K/S pair, one real per block on the same portrait.
int ell_from_ks(int nb, const int* ig, const int* jg, const double* k_di, const double* k_gg,
                const double* s_di, const double* s_gg, const int* order, SlicedEll& out);

// FoldedPreconditioner::BuildGMatrix (Code/COCR_FP/FoldedPreconditioner.cpp:68-259) as CSR in
// the reference's numbering: G^T (edge rows, tptr/tcol/tval) and G (interior-node rows,
// gptr/gcol/gval, ascending edges).
struct GradientCsr {
  int m = 0;
  std::vector<int64_t> tptr, gptr;
  std::vector<int> tcol, gcol;
  std::vector<double> tval, gval;
};
int gradient_csr(int kpar, int n_edges, int n_nodes_c, const int* nvkat, const int* nver,
                 const double* sigma3d, const int* edges, const int* ig_t, const int* jg_t,
                 const double* gg_t, const int* is_node_bound, GradientCsr& out);

// The same gradient in the device's padded layout, with the
// edges in `edge_order` (new -> old, may be null) and the interior nodes ranked by their first
// incident edge in that order.
int build_gradient(int kpar, int n_edges, int n_nodes_c, const int* nvkat, const int* nver,
                   const double* sigma3d, const int* edges, const int* ig_t, const int* jg_t,
                   const double* gg_t, const int* is_node_bound, const int* edge_order,
                   GradientEll& out);

// Structured-mesh stencil of the lowest-order edge elements. On a box of nx x ny x nz nodes the
// edges of direction d (0 x, 1 y, 2 z) at lower node (i, j, k) form boxes of
// (nx - [d==0]) x (ny - [d==1]) x (nz - [d==2]); in the grid order the rows are those boxes one
// after the other (x-edges, y-edges, z-edges), each in (k, j, i) lexicographic order. An edge
// couples with the 33 edges of its 4 cells. They are grouped in "columns" (direction d', offsets
// di, dj) holding the neighbours at dk = lo .. hi: for d = 0 and 1, 13 columns (3 parallel ones
// with dk -1..1, 4 of one cross direction with dk -1..1, 6 of z with dk -1..0); for d = 2, 21
// columns (9 parallel with dk 0, 6 + 6 cross with dk 0..1). A row's 42 coefficients are stored
// column by column and dk ascending: the 9 parallel entries complex (a, b: a + i*omega*b), the
// 24 cross entries real (a). The tables are shared by the host layout and the GPU kernel.
struct StencilCol {
  signed char dt, di, dj, lo, hi;
};
constexpr int kStcCoef = 42;              // doubles per row
constexpr int kStcMaxCols = 21;
constexpr int kStcCols[3] = {13, 13, 21};
constexpr StencilCol kStc[3][kStcMaxCols] = {
    {{0, 0, -1, -1, 1}, {0, 0, 0, -1, 1}, {0, 0, 1, -1, 1},                       // x (parallel)
     {1, 0, -1, -1, 1}, {1, 0, 0, -1, 1}, {1, 1, -1, -1, 1}, {1, 1, 0, -1, 1},   // y
     {2, 0, -1, -1, 0}, {2, 0, 0, -1, 0}, {2, 0, 1, -1, 0},                      // z
     {2, 1, -1, -1, 0}, {2, 1, 0, -1, 0}, {2, 1, 1, -1, 0}},
    {{1, -1, 0, -1, 1}, {1, 0, 0, -1, 1}, {1, 1, 0, -1, 1},                       // y (parallel)
     {0, -1, 0, -1, 1}, {0, -1, 1, -1, 1}, {0, 0, 0, -1, 1}, {0, 0, 1, -1, 1},   // x
     {2, -1, 0, -1, 0}, {2, 0, 0, -1, 0}, {2, 1, 0, -1, 0},                      // z
     {2, -1, 1, -1, 0}, {2, 0, 1, -1, 0}, {2, 1, 1, -1, 0}},
    {{2, -1, -1, 0, 0}, {2, 0, -1, 0, 0}, {2, 1, -1, 0, 0},                       // z (parallel)
     {2, -1, 0, 0, 0}, {2, 0, 0, 0, 0}, {2, 1, 0, 0, 0},
     {2, -1, 1, 0, 0}, {2, 0, 1, 0, 0}, {2, 1, 1, 0, 0},
     {0, -1, -1, 0, 1}, {0, -1, 0, 0, 1}, {0, -1, 1, 0, 1},                       // x
     {0, 0, -1, 0, 1}, {0, 0, 0, 0, 1}, {0, 0, 1, 0, 1},
     {1, -1, -1, 0, 1}, {1, 0, -1, 0, 1}, {1, 1, -1, 0, 1},                       // y
     {1, -1, 0, 0, 1}, {1, 0, 0, 0, 1}, {1, 1, 0, 0, 1}}};

// offset (in doubles) of the coefficient of column c at dk within a type-d row
constexpr int stc_slot(int d, int c, int dk) {
  int o = 0;
  for (int q = 0; q < c; ++q) {
    const int w = kStc[d][q].hi - kStc[d][q].lo + 1;
    o += w * (kStc[d][q].dt == d ? 2 : 1);
  }
  return o + (dk - kStc[d][c].lo) * (kStc[d][c].dt == d ? 2 : 1);
}

struct GridDims {
  int nx = 0, ny = 0, nz = 0;  // nodes per axis
  constexpr int nxd(int d) const { return nx - (d == 0); }
  constexpr int nyd(int d) const { return ny - (d == 1); }
  constexpr int nzd(int d) const { return nz - (d == 2); }
  constexpr int64_t count(int d) const { return (int64_t)nxd(d) * nyd(d) * nzd(d); }
  constexpr int64_t base(int d) const {
    return d == 0 ? 0 : d == 1 ? count(0) : count(0) + count(1);
  }
};

// The grid order (new -> old edge) of a structured hexahedral mesh that is one full box of
// nodes (every edge of the box present once); fails with a negative code otherwise.
int grid_order_from_mesh(int kpar, int n_nodes, const int* nver, int n_edges, const int* edges,
                         std::vector<int>& order, GridDims& dims);

// True when G (held with its edges in the grid order of `dims`) is the gradient of the whole box:
// one row per interior node in (k, j, i) order, each with the node's six edges in ascending
// reference edge order (z below, y below, x below, x above, y above, z above) and values
// +1/2, +1/2, +1/2, -1/2, -1/2, -1/2; each G^T row the edge's lower then upper node (-1/2,
// +1/2), the interior ones only. The fused grid preconditioner kernel computes this structure
// instead of reading it.
bool gradient_on_grid(const GradientEll& G, const GridDims& dims);

// Stencil coefficients (nb x kStcCoef, row-major) of a matrix held in the grid order (its row
// pack and diagonal in `E`). Fails (negative code) when a row has an entry outside its stencil
// or a complex entry in a cross position; the layout then does not apply.
int stencil_from_rows(const SlicedEll& E, const GridDims& dims, std::vector<double>& coef);

// Locality order of the edges of a structured hexahedral mesh (cells as nver[kpar][14] with the
// reference's vertex order, edges as node pairs): pencils of block x block nodes in x-y, each
// swept along z, so that the rows a GPU works on together gather from a small, cache-resident
// set of columns. Fails (negative code) when the cells are not a structured hexahedral mesh.
int order_from_mesh(int kpar, int n_nodes, const int* nver, int n_edges, const int* edges,
                    int block, std::vector<int>& order);

}  // namespace mtcocr
