/*
 * mtcocr.h — C-ABI of the MI355X COCR solver (drop-in replacement of the reference's COCR_FP).
 *
 * The reference solves one complex-symmetric system A x = b per (frequency, polarization) with
 * COCR, the folded ("DFP") preconditioner y = D_A^-1 x + G^T D_N^-1 G x and a minimal-residual
 * smoothing stop test. Entry points and the reference interfaces they replace:
 *
 *   mtcocr_pcocr_2x2_folded   PCOCR::PCOCR_2x2_Folded          Code/COCR_FP/pcocr.h:27-31,
 *                                                                  pcocr.cpp:40-126
 *   mtcocr_set_matrix_block2x2  the SLAE arrays that main() passes to it
 *                                                                  Code/COCR_FP/HarmVect3d.cpp:196-229
 *   mtcocr_set_gradient_mesh  FoldedPreconditioner::BuildGMatrix   FoldedPreconditioner.cpp:68-259
 *   mtcocr_build_gmatrix      the same, on the host, returning its CSR arrays
 *   mtcocr_solve              the request loop of PCOCR_2x2_Folded pcocr.cpp:61-123
 *   mtcocr_set_matrix_ks / mtcocr_solve_batch / mtcocr_set_rhs / mtcocr_solve_resident /
 *   mtcocr_get_solution       no single counterpart: they batch the reference's per-(f, pol)
 *                             COCR_FP processes (Code/CalcStarter/main.cpp:291-325) over one
 *                             shared sparsity, A_s = K + i*omega_s*S.
 *   mtcocr_bench_spmv, mtcocr_bench_update, mtcocr_bench_flush, mtcocr_bench_copy
 *                             no counterpart (measurement hooks: the SpMV kernel, the update
 *                             pass, the x/p pass, and a plain copy kernel that calibrates the
 *                             HBM byte counters).
 *   mtcocr_set_launch, mtcocr_get_launch, mtcocr_autotune
 *                             no counterpart (launch shapes of the GPU kernels; the reference
 *                             is a CPU OpenMP solver with no tunable launch).
 *   mtcocr_sigma_gradient     no counterpart (the conductivity gradient of an impedance misfit
 *                             after the adjoint solves; the reference computes no gradient).
 *
 * Conventions: all pointers are host pointers to caller-owned memory; indices are 0-based int32
 * (the reference decrements its 1-based files before the call, HarmVect3d.cpp:203-226); complex
 * vectors are interleaved (re, im) float64, length 2*nb; x is in/out (initial guess on entry).
 * Matrix layout = the reference's lower-triangle 2x2-block storage (Code/COCR_FP/
 * block_2x2_solver.cpp:73-105): ig[nb+1], jg[nnzL], idi[nb+1] into di, ijg[nnzL+1] into gg,
 * 1 real per block (real coupling) or 2 reals (re, im). A = L + L^T + D, no conjugation.
 *
 * Return codes: the reference's request codes (Code/COCR_FP/rci.h:23,40-42) for solver outcomes,
 * negative MTCOCR_ERR_* for argument, device and allocation errors. Nothing throws across the
 * ABI. A handle owns one HIP stream and its device memory; do not share a handle between
 * threads.
 */
#ifndef MTCOCR_H
#define MTCOCR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTCOCR_OK 0            /* REQ_OK: converged */
#define MTCOCR_X0_NULL (-1)    /* REQ_X0_NULL: ||b - A x0|| < 1e-30 */
#define MTCOCR_DIV_BY_ZERO (-2)/* REQ_DIV_BY_ZERO: (a, w) == 0 */
#define MTCOCR_MAXITER (-3)    /* REQ_MAXITER */
#define MTCOCR_ERR_ARG (-10)
#define MTCOCR_ERR_HIP (-11)
#define MTCOCR_ERR_ALLOC (-12)
#define MTCOCR_ERR_STATE (-13) /* call order: matrix / gradient / rhs missing */
#define MTCOCR_ERR_NOTCONV (-14) /* mixed precision: refinement stalled above the tolerance */

#define MTCOCR_FLAG_FP64_FINISH 1 /* mixed: refinement stalled, finished by the FP64 solver */
/* mixed: the FP64 finish met the stop test at eps/30 yet the true residual (stats.true_rel_res)
 * stayed at or above eps, the attainable accuracy of the recursively updated residual; the
 * answer is returned as the FP64 solver's would be (req MTCOCR_OK) */
#define MTCOCR_FLAG_RES_FLOOR 2

#define MTCOCR_ORDER_GRID (-1) /* mtcocr_set_mesh_order: the grid order of a box mesh */

#define MTCOCR_PREC_FP64 0  /* plain FP64 COCR, the reference's iteration and stop test */
#define MTCOCR_PREC_MIXED 1 /* complex FP32 inner COCR + FP64 iterative refinement */

typedef struct mtcocr_handle mtcocr_handle;

typedef struct mtcocr_stats {
  int req;                /* request code of this system */
  int iter;               /* COCR iterations (mixed: sum of inner iterations) */
  int outer;              /* refinement passes (FP64 path: 1) */
  int flags;              /* MTCOCR_FLAG_* */
  double residual_rel;    /* last MRS-smoothed relative residual (reference's residualRel) */
  double true_rel_res;    /* ||b - A x|| / ||b|| in FP64, recomputed on the device */
  double seconds;         /* device time of DFP setup + iterations (the reference's kit span);
                             in a batch the whole batch's time, the same for every system */
  double seconds_io;      /* host<->device copies of b and x */
} mtcocr_stats;

const char* mtcocr_error_string(int code);

/* Handle lifetime. device = HIP device ordinal. */
int mtcocr_create(int device, mtcocr_handle** out);
int mtcocr_destroy(mtcocr_handle* h);

/* One system in the reference's block layout (uploads an FP64 device copy in the solver's own
 * sliced layout). */
int mtcocr_set_matrix_block2x2(mtcocr_handle* h, int nb, const int* ig, const int* jg,
                               const int* idi, const int* ijg, const double* di,
                               const double* gg);

/* Frequency-independent pair on the same lower-triangle portrait (one real per block):
 * A(omega) = K + i*omega*S. S must vanish on blocks of real (orthogonal-edge) couplings. */
int mtcocr_set_matrix_ks(mtcocr_handle* h, int nb, const int* ig, const int* jg,
                         const double* k_di, const double* k_gg, const double* s_di,
                         const double* s_gg);

/* Discrete gradient G of the folded preconditioner from the mesh, exactly as BuildGMatrix:
 * nver[kpar*14] (0-based nodes), nvkat[kpar] (0-based materials), sigma3d[materials]
 * (a node is kept when it belongs to a cell with sigma3d > 0 and is not a boundary node),
 * edges[n_edges*2] (0-based end nodes), optional hanging-node map ig_t/jg_t/gg_t (NULL when
 * n_nodes_c covers every node), is_node_bound[n_nodes]. A row of G (one interior node) may
 * hold at most 32 entries (MTCOCR_ERR_ARG, with a message, above that): 6 on a conforming
 * hexahedral mesh, more next to hanging nodes. With a matrix set, n_edges must equal its nb;
 * setting a matrix of another size afterwards drops the gradient (the solve then returns
 * MTCOCR_ERR_STATE until it is set again). */
int mtcocr_set_gradient_mesh(mtcocr_handle* h, int kpar, int n_edges, int n_nodes_c,
                             const int* nvkat, const int* nver, const double* sigma3d,
                             const int* edges, const int* ig_t, const int* jg_t,
                             const double* gg_t, const int* is_node_bound);

/* Locality order (no counterpart in the reference, which keeps RegularMeshBuilder's
 * lexicographic edge numbering, Code/RegularMeshBuilder/main.cpp:923-1031): derives from a
 * structured hexahedral mesh (cells nver[kpar*14] in the reference's vertex order, edges[2*n]
 * 0-based node pairs, nodes < n_nodes) a device order of the edges in pencils of block x block
 * nodes swept along z, and of the interior nodes after them. The matrix, the gradient and every
 * vector are held in that order on the device; inputs and outputs keep the caller's order.
 * Must precede set_matrix_* / set_gradient_mesh (MTCOCR_ERR_STATE otherwise); block = 0 restores
 * the caller's order. block = MTCOCR_ORDER_GRID instead holds the rows in the grid order of a
 * mesh that is one full box of nx x ny x nz nodes: the x-edges, then the y- and the z-edges,
 * each set in (z, y, x) lexicographic order of its lower node; a matrix set afterwards whose
 * rows fit the edge-element stencil (33 couplings at fixed offsets) also gets the structured
 * stencil SpMV layout (mtcocr_set_launch layout 3). Per-row arithmetic is unchanged (each row keeps its entry order), dot
 * products regroup. Returns MTCOCR_ERR_ARG, keeping the caller's order, for a mesh that is not
 * structured hexahedral. mtcocr_pcocr_2x2_folded and the COCR_FP executable apply the grid
 * order where the mesh is a box, else block 16. */
int mtcocr_set_mesh_order(mtcocr_handle* h, int kpar, int n_nodes, const int* nver, int n_edges,
                          const int* edges, int block);

/* The same order on the host, without a device: order[i] = caller's edge index of device row i
 * (n_edges entries). */
int mtcocr_mesh_order(int kpar, int n_nodes, const int* nver, int n_edges, const int* edges,
                      int block, int* order);

/* FoldedPreconditioner::BuildGMatrix (Code/COCR_FP/FoldedPreconditioner.cpp:68-259) on the
 * host, without a device: the discrete gradient as the reference's CSR arrays, 0-based and in
 * its node numbering (interior nodes of conductive cells, ascending), G^T (edge rows:
 * iptr_gt[n_edges_c+1], jptr_gt[nnz], gg_gt[nnz]) and G (node rows, ascending edges:
 * iptr_g[m+1], jptr_g[nnz], gg_g[nnz]). Arguments as mtcocr_set_gradient_mesh; hanging nodes
 * (numbers >= n_nodes_c) spread through ig_t/jg_t/gg_t, indexed by their global number as the
 * reference's HarmVect3d.cpp:255-301 builds them. Call once with the six output arrays NULL
 * to get m and nnz, then with arrays of those sizes. mtcocr_set_gradient_mesh builds the same
 * matrix and rejects (MTCOCR_ERR_ARG) one with more than 32 entries in a row of G. */
int mtcocr_build_gmatrix(int kpar, int n_edges_c, int n_nodes_c, const int* nvkat,
                         const int* nver, const double* sigma3d, const int* edges,
                         const int* ig_t, const int* jg_t, const double* gg_t,
                         const int* is_node_bound, int* m, int64_t* nnz, int* iptr_gt,
                         int* jptr_gt, double* gg_gt, int* iptr_g, int* jptr_g, double* gg_g);

/* Solve one system set by mtcocr_set_matrix_block2x2. pr = b (2*nb), x in/out. */
int mtcocr_solve(mtcocr_handle* h, const double* pr, double* x, double eps, int maxiter,
                 int precision, mtcocr_stats* stats);

/* Batched solve of nsys (1..64) systems A(omega_s) x_s = b_s sharing the matrix set by
 * mtcocr_set_matrix_ks, or, with omega == NULL, nsys right-hand sides of the block2x2 matrix
 * (for example both polarizations of one frequency).
 * pr and x hold nsys consecutive vectors of 2*nb doubles; stats has nsys entries.
 * Systems retire independently. Returns 0 when every system returned MTCOCR_OK, otherwise the
 * first non-zero per-system code. */
int mtcocr_solve_batch(mtcocr_handle* h, int nsys, const double* omega, const double* pr,
                       double* x, double eps, int maxiter, int precision, mtcocr_stats* stats);

/* Resident-input form of mtcocr_solve_batch, for callers that time the device work alone:
 * set_rhs copies b (and x0) to the device, solve_resident runs the solver on device-resident
 * data and waits for it, get_solution copies x back. */
int mtcocr_set_rhs(mtcocr_handle* h, int nsys, const double* omega, const double* pr,
                   const double* x0);
int mtcocr_solve_resident(mtcocr_handle* h, double eps, int maxiter, int precision,
                          mtcocr_stats* stats);
int mtcocr_get_solution(mtcocr_handle* h, double* x);

/* One-shot form with the argument list of PCOCR::PCOCR_2x2_Folded minus y_omp (the reference's
 * per-thread scatter buffer) and with nver/edges as flat arrays. Runs on device 0 (or the device
 * named by MTCOCR_DEVICE). */
int mtcocr_pcocr_2x2_folded(int n, const int* ig, const int* jg, const int* idi, const int* ijg,
                            const double* di, const double* gg, const double* pr, double* x,
                            double eps, int maxiter, int kpar, int n_edges_c, int n_nodes_c,
                            const double* xyz, const int* nvkat, const int* nver,
                            const double* sigma3d, const int* edges, const int* ig_t,
                            const int* jg_t, const double* gg_t, const int* is_node_bound,
                            mtcocr_stats* stats);

/* Measurement hook: runs the SpMV kernel `reps` times on the handle's stream (precision 0: FP64
 * vectors, 1: FP32 vectors; the matrix is FP64 in both) over the nsys systems of the current
 * batch and returns the average kernel time from HIP events and the algorithmic bytes moved per
 * launch (matrix once, x and y once per system). */
int mtcocr_bench_spmv(mtcocr_handle* h, int precision, int nsys, int reps, double* avg_ms,
                      double* bytes_per_launch);

/* Launch shapes: workgroup caps of the grid-stride SpMV, edge-streaming and node kernels
 * (multiples of 8 in [8, 131072]; 0 keeps a value) and the SpMV layout (-1 keeps it):
 *   0 sliced ELL (any batch width);
 *   1 row pack, one row per group of lanes (batches of 3..64 systems);
 *   2 row blocks: the row pack over 16-row blocks whose x values are staged in LDS (batches of
 *     3..16 systems, when no block has more than 192 distinct columns); where it does not
 *     apply, the row pack runs;
 *   3 structured stencil: 42 coefficients per row at fixed stencil positions, each wave
 *     sweeping a column of rows along z (batches of 9..16 systems, matrix set in the grid order
 *     and fitting the stencil); where it does not apply, the row blocks run;
 *   4 the same stencil with x staged in LDS: each workgroup sweeps a 4 x 4 tile of rows and
 *     copies the tile's x neighbourhood of every level into LDS once (same conditions; each
 *     row's result equals layout 3's bit for bit).
 * The row layouts need the batch vectors under 4 GiB (nb x width x 16 bytes in FP64); above
 * that the sliced ELL runs. set_launch applies to both precisions and every batch width;
 * get_launch reports those of `precision` at the current batch width, the layout as the one
 * that runs. Caps change only the reduction grouping of the dot products and the layout only
 * the summation order inside a row (results are deterministic for a fixed shape). cap_vec
 * also sets the fused update pass's cap, which the autotuner tunes apart. Defaults: caps
 * 8192 / 65536 / 2048 (update pass 131072), sliced ELL for 1..2 systems, the LDS-staged
 * stencil (layout 4; layout 3 is the wave-task stencil) for 9..16 where it applies, else row
 * blocks for 3..16, row pack for 17..64. */
int mtcocr_set_launch(mtcocr_handle* h, int cap_spmv, int cap_vec, int cap_node,
                      int spmv_layout);
int mtcocr_get_launch(mtcocr_handle* h, int precision, int* cap_spmv, int* cap_vec,
                      int* cap_node, int* spmv_layout);

/* Sweep length of the stencil SpMVs (layouts 3 and 4): rows of one edge column that a task
 * sweeps along z (at least 1; default: whole columns; longer sweeps reload fewer levels,
 * shorter ones make more tasks and keep the concurrently swept region of x within the L2
 * cache). set applies to every width and precision; get reports the value at the current
 * batch width. No counterpart in the reference. */
int mtcocr_set_stencil_sweep(mtcocr_handle* h, int kc);
int mtcocr_get_stencil_sweep(mtcocr_handle* h, int precision, int* kc);

/* Task order of the LDS-staged stencil SpMV (layout 4): bands of `band` tile columns (4 rows
 * each in x), each band's tiles row by row, so that the workgroups of one XCD work on a patch
 * whose tiles share their x halos within one L2; 0 (the default) orders all tile columns as
 * one band. Only the order changes: every product is the same, the dot products' partial
 * sums regroup. set applies to every width and precision; get reports the value at the
 * current batch width. The autotuner searches it. No counterpart in the reference. */
int mtcocr_set_stencil_band(mtcocr_handle* h, int band);
int mtcocr_get_stencil_band(mtcocr_handle* h, int precision, int* band);

/* The FP64 solver's x/p period: x and p are updated every `period` iterations (8 or 16) from
 * a ring of 2 * period s vectors (the reference updates them every iteration,
 * pcocr_rci.cpp:140-142,164; the answer is the same bit for bit). 16 moves fewer bytes per
 * iteration but holds 16 more batch vectors on the device; 0 (the default) takes 16 when, at
 * the first FP64 solve of a batch shape, they leave 8 GiB and the compaction buffers free,
 * else 8. set 8 frees the extra vectors. get reports the period in use. No counterpart in the
 * reference. */
int mtcocr_set_xp_period(mtcocr_handle* h, int period);
int mtcocr_get_xp_period(mtcocr_handle* h, int* period);

/* Autotuner: for the batch currently on the handle (set by mtcocr_set_rhs or a previous
 * mtcocr_solve_batch; needs the matrix and the gradient mesh) times the SpMV (all layouts),
 * node, edge and update-pass kernels at caps 512..131072, `reps` launches each, in
 * `precision`, and keeps the fastest choice of each family on the handle for this precision
 * and batch width (padded system count). FP64 batches wider than 16 systems run their
 * iterations as groups of 16: they are tuned at width 16, which the outputs then report, while
 * mtcocr_get_launch reports the width of the batch itself (used by the kernel-level hooks).
 * In FP64 the widths that compaction narrows a batch to (half, a quarter, ... one system) are
 * tuned too, on the batch's first systems. Outputs (any may be NULL): the chosen caps and layout, and the SpMV time in ms at the chosen
 * shape. Writes only work buffers: resident right-hand sides and solutions stay. */
int mtcocr_autotune(mtcocr_handle* h, int precision, int reps, int* cap_spmv, int* cap_vec,
                    int* cap_node, int* spmv_layout, double* spmv_ms);

/* Kernel-level timing hook, no counterpart in the reference (its per-iteration vector updates,
 * pcocr_rci.cpp:135-154 and MRS.cpp:85-153, run as separate loops): average time (ms) of the
 * fused update pass (k_xr_fused: the MRS blend, r -= alpha a, s -= alpha w and their partials;
 * in the FP32-vector inner solve of the mixed solver also p = s + beta p, x += alpha p and the
 * smoothed x) over the systems of the current batch, `reps` launches on filled work vectors,
 * and the algorithmic bytes per launch (each stream read or written once per system and edge:
 * 8 streams in FP64, whose x and p are updated by k_flush_xp every eighth iteration; in FP64
 * on a box mesh in the grid order with the fused grid preconditioner, the pass that forms w
 * from the node vector help: 7 streams plus help once per interior node and system and the
 * diagonal once per edge; 14 streams with the FP32 vectors). */
int mtcocr_bench_update(mtcocr_handle* h, int precision, int reps, double* avg_ms,
                        double* bytes_per_launch);

/* Kernel-level timing hook, no counterpart in the reference (pcocr_rci.cpp:140-142,164 update
 * p and x every iteration): average time (ms) of the FP64 solver's x/p pass (k_flush_xp: eight
 * iterations of p = s + beta p, x += alpha p from eight s vectors of its ring, run every eighth
 * iteration) over the systems of the current batch, `reps` launches, and its algorithmic bytes
 * per launch (x, p and eight s read, x and p written: 12 streams per system and edge). */
int mtcocr_bench_flush(mtcocr_handle* h, int reps, double* avg_ms, double* bytes_per_launch);

/* Calibration hook: average time (ms) of a plain device-to-device copy kernel (k_copy, 16-byte
 * elements) over nbytes (a multiple of 16), `reps` launches after one warm-up. Its HBM byte
 * counters (read + write of nbytes each) calibrate the counters of the solver's kernels. */
int mtcocr_bench_copy(mtcocr_handle* h, int64_t nbytes, int reps, double* avg_ms);

/* Kernel-level hook for parity checks: y = op(x) on the current matrix, op 0 = SpMV (A x),
 * op 1 = folded preconditioner (D_A^-1 x + G^T D_N^-1 G x), in precision 0 (FP64) or 1 (FP32
 * storage, FP64 in/out). Uses system 0 of the current batch (omega = 1 for a block2x2 matrix).
 * Counterparts: Block_2x2_solver::Mult_MV_block_2x2 (block_2x2_solver.cpp:73-173) and
 * FoldedPreconditioner::ApplyPreconditionerComplex (FoldedPreconditioner.cpp:578-629). */
int mtcocr_apply(mtcocr_handle* h, int op, int precision, const double* x, double* y);

/* Allocation-free form of mtcocr_apply, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. x and y are DEVICE pointers to one complex vector each (2*nb
 * interleaved doubles, FP64) in the caller's row order, so the handle must hold no device order
 * (mtcocr_set_mesh_order not applied: MTCOCR_ERR_STATE otherwise). omega: the frequency of a
 * K/S matrix (A = K + i*omega*S); 1 for a block2x2 matrix. op as mtcocr_apply (0 SpMV,
 * 1 preconditioner). The work goes to `stream` (a hipStream_t; NULL: the handle's stream) and
 * is not waited for. It calls no allocating routine: the D_N^-1 diagonal, the node vector of
 * the preconditioner and the SpMV's partial sums live in a 32 MiB scratch array that every
 * GPU the library runs on holds for the life of the process, which limits the preconditioner to
 * meshes of about one million interior nodes (MTCOCR_ERR_ARG beyond, with a message). All
 * calls on one GPU share that array: calls that overlap on one GPU (other streams, other
 * threads, other handles) race and must be serialised by the caller. The result equals
 * mtcocr_apply's on the same handle without a batch, bit for bit. */
int mtcocr_apply_noalloc(mtcocr_handle* h, int op, double omega, const double* x, double* y,
                         void* stream);

/* Matrix statistics of the device layout: complex / real non-zeros of the full matrix (explicit
 * zeros dropped), number of interior nodes m of G, device bytes held by the handle. */
int mtcocr_matrix_info(mtcocr_handle* h, int64_t* nnz_complex, int64_t* nnz_real, int64_t* m,
                       int64_t* device_bytes);

/* Conductivity gradient contraction of the adjoint, on the device. No counterpart in the
 * reference (it computes no gradient); it completes dJ/dsigma after the adjoint solves:
 *   grad[c] = sum_s Re( coef_s * u_s^T M_c v_s )
 * over the cells c of a regular grid of nxn x nyn x nzn nodes (coordinates x, y, z; cells with
 * x fastest), M_c the cell's 12x12 edge mass matrix without sigma (the same matrix as
 * mtmodel_element_matrices' c: dA/dsigma_c = i*omega M_c), u_s and v_s complex interleaved
 * edge vectors of system s in the mesh writer's edge order (host arrays of nsys x 2*n_edges),
 * coef_s complex interleaved (2*nsys). For the impedance misfit, u_s = the adjoint solution,
 * v_s = the total field (solution plus normal-field DOFs) and coef_s = -2i*omega_s. One thread
 * per cell sums its systems in order (deterministic). Needs no matrix on the handle. */
int mtcocr_sigma_gradient(mtcocr_handle* h, int nxn, int nyn, int nzn, const double* x,
                          const double* y, const double* z, int nsys, const double* coef,
                          const double* u, const double* v, double* grad);

#ifdef __cplusplus
}
#endif

#endif /* MTCOCR_H */
