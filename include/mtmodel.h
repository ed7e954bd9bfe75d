/*
 * mtmodel.h — C-ABI of the regular-mesh model layer: the producer of the COCR
 * solver's input on a regular hexahedral mesh.
 *
 * It restates, for regular tensor meshes, what three reference stages write to
 * disk before COCR_FP runs:
 *   - mesh topology and numbering        Code/RegularMeshBuilder/main.cpp:923-1031,1347-1797
 *   - the 1-D layered normal field A0(z) Code/Harm1D/divgrad_1d.cpp:147-200,
 *                                        mesh_1d_mtz.cpp:95-187, local_matrix_1d.cpp:45-103,
 *                                        global_slae_1d_harm_prof.cpp:74-138
 *   - edge-element assembly + Dirichlet   Code/BuildMatrix/t_global_slae.cpp:80-172,
 *                                        T_Portrait.cpp:47-188, T_Brick.cpp:1361-1384,
 *                                        bound_cond_vec_harm.cpp:260-307
 *
 * Conventions (all arrays are caller-owned, plain C types):
 *   nxn, nyn, nzn  node counts per axis (cells = n-1 per axis); x[nxn], y[nyn], z[nzn] ascending.
 *   Nodes are numbered x fastest, then y, then z (0-based here, 1-based in files).
 *   Edges are numbered by the lexicographic order of (first node, second node): for each node
 *   in order, its +x edge, then +y edge, then +z edge when they exist.
 *   Cells are numbered x fastest, then y, then z; a cell's 12 edges follow the reference's
 *   local order (0-3 x-edges, 4-7 y-edges, 8-11 z-edges).
 *   "Block lower-triangle" storage is the reference's SLAE file layout, 0-based:
 *   ig[nb+1], jg[nnzL] (columns < row, ascending), idi[nb+1] (offsets into di),
 *   ijg[nnzL+1] (offsets into gg); a block holds 1 real (orthogonal edges) or 2 reals
 *   (re, im: parallel edges and all diagonals).
 *
 * Return value of every int function: 0 on success, negative MTMODEL_ERR_* on failure.
 */
#ifndef MTMODEL_H
#define MTMODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTMODEL_ERR_ARG   (-10)
#define MTMODEL_ERR_SIZE  (-11)
#define MTMODEL_ERR_ALLOC (-12)
#define MTMODEL_ERR_SOLVE (-13)

/* Number of edges (complex unknowns) of the mesh. */
int64_t mtmodel_num_edges(int nxn, int nyn, int nzn);

/* File-format topology arrays, 1-based like the reference's binary files:
 *   nver[kpar*14]   (nver.dat: 8 corner nodes then 6 zeros)      RegularMeshBuilder/main.cpp:1389-1447
 *   ed[kpar*25]     (edges.dat: 12 edge ids, 12 zeros, 1)          main.cpp:1766-1797
 *   nodes_for_edges[nb*2] (nodesforedges.dat)                      main.cpp:1738-1763
 *   l13d[kt1]       (L13d.dat: boundary nodes in node order)        main.cpp:1706-1735
 * Any pointer may be NULL to skip that array. kt1 = number of boundary nodes. */
int mtmodel_topology(int nxn, int nyn, int nzn, int* nver, int* ed, int* nodes_for_edges,
                     int* l13d, int64_t* kt1);

/* Number of strictly-lower blocks of the edge portrait (T_Portrait::Gen_Portrait). */
int mtmodel_portrait_size(int nxn, int nyn, int nzn, int64_t* nnz_lower);

/* 0-based portrait and block-size prefix arrays (T_Portrait::Gen_Portrait + Gen_idi_ijg). */
int mtmodel_portrait(int nxn, int nyn, int nzn, int* ig, int* jg, int* idi, int* ijg);

/* 1-D layered normal field (Harm1D). Layers are given bottom-first: layer_top[i] is the top of
 * layer i (layer_top[n_layers-1] must be 0, the earth surface); layer_sigma[i] its conductivity.
 * The deepest layer extends to z = -1e7 m. nu is the frequency in Hz.
 * Writes up to max_points values of z (ascending), usin = Re A0, ucos = Im A0, each rounded
 * through the reference's "%25.13e" text files; *n_points receives the count. */
int mtmodel_harm1d(int n_layers, const double* layer_top, const double* layer_sigma, double nu,
                   int max_points, int* n_points, double* z1d, double* usin, double* ucos);

/* System of one (frequency, polarization) in the reference's SLAE file layout, homogeneous
 * Dirichlet conditions applied on boundary edges. Per-cell sigma, sigma0 (normal/background)
 * and absolute mu. alpha = 1 for the X polarization (A0 along x), 0 for Y.
 * Outputs: di[idi[nb]], gg[ijg[nnzL]], pr[2*nb]. */
int mtmodel_assemble(int nxn, int nyn, int nzn, const double* x, const double* y, const double* z,
                     const double* cell_sigma, const double* cell_sigma0, const double* cell_mu,
                     double nu, int alpha, int n1d, const double* z1d, const double* usin,
                     const double* ucos, const int* ig, const int* jg, const int* idi,
                     const int* ijg, double* di, double* gg, double* pr);

/* Right-hand side alone (same arithmetic as mtmodel_assemble): pr[2*nb]. */
int mtmodel_assemble_rhs(int nxn, int nyn, int nzn, const double* x, const double* y,
                         const double* z, const double* cell_sigma, const double* cell_sigma0,
                         double nu, int alpha, int n1d, const double* z1d, const double* usin,
                         const double* ucos, double* pr);

/* Frequency-independent split A(omega) = K + i*omega*S on the same portrait, one real per block
 * (k_di[nb], k_gg[nnzL], s_di[nb], s_gg[nnzL]); Dirichlet applied (K diagonal 1, S diagonal 0,
 * boundary rows and columns zero). K = sum of curl-curl/mu, S = sum of sigma*mass. */
int mtmodel_assemble_ks(int nxn, int nyn, int nzn, const double* x, const double* y,
                        const double* z, const double* cell_sigma, const double* cell_mu,
                        const int* ig, const int* jg, double* k_di, double* k_gg, double* s_di,
                        double* s_gg);

/* Element matrices of a brick hx*hy*hz in the reference's local edge order, row-major 12x12:
 * b = curl-curl (no 1/mu), c = mass (no sigma). Used by tests and by the adjoint. */
int mtmodel_element_matrices(double hx, double hy, double hz, double* b, double* c);

/* Edge DOFs (complex interleaved, a0[2*nb]) of the 1-D normal field of polarization alpha, the
 * values the right-hand side is built from: (A0(z_e) (alpha, 1 - alpha, 0) . t_e) h_e / 2 at the
 * height of horizontal edges, 0 on vertical edges (Code/BuildMatrix/T_Brick.cpp:1361-1495
 * evaluates the same normal field per element). Used by the adjoint: the total field is the
 * solution plus these DOFs. */
int mtmodel_normal_dofs(int nxn, int nyn, int nzn, const double* x, const double* y,
                        const double* z, int alpha, int n1d, const double* z1d,
                        const double* usin, const double* ucos, double* a0);

/* Per cell c: out[2c], out[2c+1] = u_c^T M_c v_c (complex, unconjugated) with M_c the cell's
 * 12x12 mass matrix (the sigma-derivative of the system matrix up to i*omega) and u, v
 * complex interleaved edge vectors. No counterpart in the reference (it has no gradient); used
 * by the adjoint sensitivity dJ/dsigma_c = 2 Re[-i*omega lambda_c^T M_c (x + a0)_c]. */
int mtmodel_cell_mass_forms(int nxn, int nyn, int nzn, const double* x, const double* y,
                            const double* z, const double* u, const double* v, double* out);

#ifdef __cplusplus
}
#endif

#endif /* MTMODEL_H */
