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

// Locality order of the edges of a structured hexahedral mesh (cells as nver[kpar][14] with the
// reference's vertex order, edges as node pairs): pencils of block x block nodes in x-y, each
// swept along z, so that the rows a GPU works on together gather from a small, cache-resident
// set of columns. Fails (negative code) when the cells are not a structured hexahedral mesh.
int order_from_mesh(int kpar, int n_nodes, const int* nver, int n_edges, const int* edges,
                    int block, std::vector<int>& order);

}  // namespace mtcocr
