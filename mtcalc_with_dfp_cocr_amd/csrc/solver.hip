// MI355X COCR solver for the complex-symmetric edge-element system of the MT forward problem.
//
// Method (the reference's Code/COCR_FP, restated for the GPU):
//   COCR request order   Code/COCR_FP/pcocr_rci.cpp:50-181 (stages 1-8)
//   stop test            Code/COCR_FP/pcocr.cpp:91-96, rci.cpp:66-90 (MRS residual / ||b||)
//   MRS smoothing        Code/COCR_FP/MRS.cpp:85-153 (real-valued, over the 2*nb reals)
//   preconditioner       Code/COCR_FP/FoldedPreconditioner.cpp:473-629 (diag(A)^-1 +
//                        G^T diag(G A G^T)^-1 G)
//   SpMV                 Code/COCR_FP/block_2x2_solver.cpp:73-173 (A = L + L^T + D)
//   dot products         Code/COCR_FP/base_solver.cpp:382-399 (unconjugated)
//
// Device design (one GPU, HIP streams and graphs):
//   * S systems share one sparsity; vectors are laid out [row][system] so that a gather of
//     x[col] for all systems is one contiguous segment. A single system is S = 1.
//   * The matrix is a sliced ELL of 64-row slices (host_prep.h), complex entries a + i*omega*b,
//     read once per row group and reused by every system of the batch.
//   * Every reduction is deterministic: fixed grid, per-block partials in a fixed order, then a
//     one-block "scalar" kernel that also carries the solver's scalar logic (alpha, beta, MRS
//     weight, stop test, iteration count). No atomics, no inter-block fences.
//   * Per-system status flags retire systems independently; the host polls a running count
//     every chunk of iterations, and a chunk is replayed from a captured hipGraph.
//   * Mixed precision: complex FP32 inner COCR (SpMV, preconditioner, updates), FP64 outer
//     iterative refinement (true residual, update, final residual).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "host_prep.h"
#include "mtcocr.h"
#include "mtmodel.h"

namespace mtcocr {

// ------------------------------------------------------------------------------------------
// small helpers

constexpr int kBlock = 256;      // threads per block of the streaming kernels
constexpr int kMaxBlocks = 131072; // largest grid of a reduction kernel (size of the partials)
constexpr int kScalarThreads = 1024;
constexpr int kXcd = 8;          // MI355X: 8 XCDs, workgroups dispatched round-robin

enum : int { ST_RUNNING = 1, ST_INACTIVE = 2 };
enum : int { MRS_SKIP = 0, MRS_COPY = 1, MRS_BLEND = 2, MRS_NONE = 3 };

template <class T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

// Streaming loads and stores with the nontemporal hint, for vectors that a kernel touches once
// (they are far larger than L2 and the MALL, so caching them only evicts the gathered data).
template <class T2>
__device__ inline T2 ld_nt(const T2* p) {
  using T = decltype(T2::x);
  using V = T __attribute__((ext_vector_type(2)));
  const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
  T2 r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
template <class T2>
__device__ inline void st_nt(T2* p, T2 o) {
  using T = decltype(T2::x);
  using V = T __attribute__((ext_vector_type(2)));
  V v;
  v.x = o.x;
  v.y = o.y;
  __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
}

// The FP64 solver's x/p recurrence runs every P iterations (P = 8 or 16 per handle, by the
// device memory free for its ring) from a ring of 2P s buffers (s_{k-P}..s_{k-1} are read
// while the update pass has written s_k), and its graphs hold 2P iterations, so that each
// position of the ring is a fixed buffer in them.
constexpr int kXpMax = 16;
constexpr int kRingMax = 2 * kXpMax;

// Per-system scalar state (double precision in every mode).
struct SysScalars {
  double zs_re, zs_im, zs1_re, zs1_im;
  double aw_re, aw_im, al_re, al_im, be_re, be_im;
  double r_old, res, rel, w, eps;
  double scale, prev_rel, b_norm, best_rel;
  int mode, status, iter, total_iter;
  int outer_status, outer, last_improve, stall_window;
  int fallback, fallback_iter, rollback, halt;
  int last, pad1, pad2, pad3;  // last: the iteration whose x is the answer (buffer last % 2)
  // alpha (re, im) and beta (re, im) of iteration k at position (k - 1) % 32: the FP64 solver
  // applies the x/p recurrence eight or sixteen iterations at a time from these
  double hab[kRingMax][4];
};

template <class T>
struct EllView {
  const int *offc, *nc, *offr, *nr, *colc, *colr;
  const T *ac, *bc, *ar;
  int nb;
};

// row pack: row-contiguous off-diagonal entries, complex ones first (see host_prep.h), and the
// diagonal as da + i*omega*db; FP64 values
struct RowView {
  const int *rp, *rb, *col;
  const double *a, *b, *da, *db;
  int nb;
};

// The row pack seen through 16-row blocks: block k's distinct columns are
// halo[hoff[k] .. hoff[k+1]); lcol (parallel to the row pack's columns) and ldiag give each
// entry's and each row's position in its block's halo.
struct BlkView {
  const int *rp, *rb, *hoff, *halo;
  const unsigned short *lcol, *ldiag;
  const double *a, *b, *da, *db;
  int nb, nblk;
  long long nhalo, nnz, nnzb;  // lengths of halo, a (and lcol), b
};

constexpr int kRowBlock = 16;       // rows per block of the halo-staged SpMV
// Longest row of G (entries of one interior node) the preconditioner set-up takes. A node of a
// conforming hexahedral mesh has at most 6 edges; hanging nodes of a non-conforming mesh add
// the edges of their T-matrix rows (a 2:1 face adds up to 12 more).
constexpr int kMaxGRow = 32;
constexpr int kHaloStage = 192;     // largest block halo it stages (columns)

template <class T>
struct GView {
  const int *gcol, *gtcol;
  const T *gval, *gtval;
  int m, n, wg, wgt;
};

// complex division with scaling (Smith), used for alpha, beta and reciprocals
__host__ __device__ inline void cdiv(double ar, double ai, double br, double bi, double& qr,
                                     double& qi) {
#pragma clang fp contract(off)
  if (fabs(br) >= fabs(bi)) {
    const double r = bi / br, d = br + bi * r;
    qr = (ar + ai * r) / d;
    qi = (ai - ar * r) / d;
  } else {
    const double r = br / bi, d = bi + br * r;
    qr = (ar * r + ai) / d;
    qi = (ai * r - ar) / d;
  }
}

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Grid of a grid-stride kernel over n elements: one workgroup per 256-element tile up to `cap`
// workgroups (the launch shape the autotuner picks), a multiple of the XCD count.
inline int grid_for(long long n, int cap) {
  long long b = (n + kBlock - 1) / kBlock;
  if (cap > kMaxBlocks) cap = kMaxBlocks;
  if (b > cap) b = cap;
  if (b < kXcd) b = kXcd;
  return round_up(int(b), kXcd);
}

// XCD-aware tile walk: workgroup b runs on XCD b % 8; each XCD owns a contiguous range of
// 256-element tiles so neighbouring rows (and their x gathers) share one L2.
struct TileWalk {
  long long first, last, step;
  __device__ TileWalk(long long ntiles) {
    const int nblk = gridDim.x;
    const int xcd = blockIdx.x % kXcd;
    const int lb = blockIdx.x / kXcd;
    const int per = nblk / kXcd;
    const long long lo = ntiles * xcd / kXcd;
    const long long hi = ntiles * (xcd + 1) / kXcd;
    first = lo + lb;
    last = hi;
    step = per;
  }
};

// Block reduction of NV values per system; lanes with equal threadIdx % SP share a system.
// Writes partials[blockIdx.x * (SP*NV) + s*NV + k].
template <int SP, int NV>
__device__ inline void block_partials(double (&v)[NV], double* partials) {
  for (int off = SP; off < 64; off <<= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, 64);
  __shared__ double sh[kBlock / 64][64 * NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < SP)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[wave][lane * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x < SP * NV) {
    double s = 0.0;
    for (int wv = 0; wv < kBlock / 64; ++wv) s += sh[wv][threadIdx.x];
    partials[(size_t)blockIdx.x * (SP * NV) + threadIdx.x] = s;
  }
}

// block_partials for kernels with 16-lane rows of 8 systems x 2 lanes (lane = 16 g + 2 s + c),
// once the two lanes of a system hold the same sums: lanes with equal (lane % 16) / 2 share a
// system. Writes partials[blockIdx.x * (SP*NV) + s*NV + k], SP = 8.
template <int SP, int NV>
__device__ inline void block_partials_split(double (&v)[NV], double* partials) {
  static_assert(SP == 8, "8 systems on 16-lane rows");
  for (int off = 16; off < 64; off <<= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, 64);
  __shared__ double sh[kBlock / 64][SP * NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < 16 && (lane & 1) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[wave][(lane >> 1) * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x < SP * NV) {
    double s = 0.0;
    for (int wv = 0; wv < kBlock / 64; ++wv) s += sh[wv][threadIdx.x];
    partials[(size_t)blockIdx.x * (SP * NV) + threadIdx.x] = s;
  }
}

// One-block reduction of nblk x P partials into tot[P] (fixed order): G = kScalarThreads / P
// threads per column each sum a strided subset, then a fixed tree over the G sums (the part
// of G above its largest power of two folded in first).
template <int P>
__device__ inline void reduce_partials(const double* part, int nblk, double* tot) {
  constexpr int G = kScalarThreads / P;
  constexpr int G2 = G >= 512 ? 512 : G >= 256 ? 256 : G >= 128 ? 128 : G >= 64 ? 64
                   : G >= 32 ? 32 : G >= 16 ? 16 : G >= 8 ? 8 : G >= 4 ? 4 : G >= 2 ? 2 : 1;
  __shared__ double tmp[kScalarThreads];
  const int t = threadIdx.x;
  const int q = t % P, j = t / P;
  double acc = 0.0;
  if (j < G) {
    // 16 loads of a thread in flight together, rows past nblk read as 0; one accumulator in
    // row order
    constexpr int U = 16;
    for (int b0 = j; b0 < nblk; b0 += U * G) {
      double v[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int b = b0 + k * G;
        v[k] = b < nblk ? part[(size_t)b * P + q] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < U; ++k) acc += v[k];
    }
  }
  tmp[t] = acc;
  __syncthreads();
  if (G2 < G) {
    if (j < G - G2) tmp[t] += tmp[t + G2 * P];
    __syncthreads();
  }
  for (int hw = G2 / 2; hw >= 1; hw /= 2) {
    if (j < hw) tmp[t] += tmp[t + hw * P];
    __syncthreads();
  }
  if (t < P) tot[t] = tmp[t];
  __syncthreads();
}

__device__ inline void count_running(const SysScalars* sc, int nsp, int* nrun) {
  // called by thread 0 after the per-system updates
  int c = 0;
  for (int s = 0; s < nsp; ++s) c += (sc[s].status == ST_RUNNING);
  *nrun = c;
}

// ------------------------------------------------------------------------------------------
// element-wise complex helpers: z = x + a*y written exactly as Base_solver::Cmplx_axpy.
// Contraction is off so that every kernel that evaluates the same update gets the same bits:
// the preconditioner's node kernel recomputes a = z + beta*a from z and the old a instead of
// reading the edge kernel's stored result, and in the air G*a is a tiny difference that D_N^-1
// amplifies, so a one-ulp disagreement between the two copies of a (FP32) breaks COCR.

template <class T2>
__device__ inline T2 caxpy(T2 x, T2 y, double ar, double ai) {
#pragma clang fp contract(off)
  using T = decltype(x.x);
  const T a_re = T(ar), a_im = T(ai);
  T2 r;
  r.x = x.x + (y.x * a_re - y.y * a_im);
  r.y = x.y + (y.x * a_im + y.y * a_re);
  return r;
}

// ------------------------------------------------------------------------------------------
// kernels

// plain streaming copy (16-byte elements): the calibration kernel for the HBM byte counters
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src,
                                              double2* __restrict__ dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = i; k < n; k += stride) dst[k] = src[k];
}

template <class T2>
__global__ void k_fill(T2* v, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = i; k < n; k += stride) {
    T2 z;
    z.x = 0;
    z.y = 0;
    v[k] = z;
  }
}

// Conductivity gradient contraction of the adjoint: grad[c] = sum_s Re(coef_s * u_s^T M_c v_s)
// for the cells of a regular grid (x fastest), one thread per cell. M_c is the lowest-order
// edge mass matrix of the brick hx*hy*hz in the mesh writer's local edge order (edges 0-3 along
// x, 4-7 along y, 8-11 along z); only edges of one direction couple. Its entries are formed as
// the model layer's element integrals form them: for edges i, j of direction d, sitting at
// signs s_i, s_j in the two other coordinates of the reference cube [-1,1]^3,
//   M_ij = (1/(2 h_d))^2 * prod_k I_k * hx hy hz / 8,  I_d = 2,  I_k = 2 + (2/3) s_ik s_jk,
// the integral of the product of the two basis functions (edge DOF = tangential field * h/2).
// ce holds the 12 edges of every cell; u and v are [system][edge] complex.
__global__ __launch_bounds__(256) void k_sigma_grad(
    const int* __restrict__ ce, const double* __restrict__ xn, const double* __restrict__ yn,
    const double* __restrict__ zn, int ncx, int ncy, long long ncells, int nsys,
    const double* __restrict__ coef, const double2* __restrict__ u,
    const double2* __restrict__ v, long long nb, double* __restrict__ grad) {
#pragma clang fp contract(off)
  // position signs of the local edges in the two coordinates other than their direction
  constexpr signed char kSign[12][3] = {
      {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},  {-1, 0, -1}, {-1, 0, 1},
      {1, 0, -1},  {1, 0, 1},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += stride) {
    const int cx = int(c % ncx), cy = int((c / ncx) % ncy), cz = int(c / ((long long)ncx * ncy));
    const double h[3] = {xn[cx + 1] - xn[cx], yn[cy + 1] - yn[cy], zn[cz + 1] - zn[cz]};
    const double detj = h[0] * h[1] * h[2] / 8.0;
    double m[3][4][4];
    for (int d = 0; d < 3; ++d) {
      const double base = 0.5 / h[d];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          double t = base * base;
          for (int k = 0; k < 3; ++k)
            t *= 2.0 + (2.0 / 3.0) * double(kSign[4 * d + i][k]) * double(kSign[4 * d + j][k]);
          m[d][i][j] = t * detj;
        }
    }
    int e[12];
    for (int k = 0; k < 12; ++k) e[k] = ce[c * 12 + k];
    double g = 0.0;
    for (int s = 0; s < nsys; ++s) {
      const double2* us = u + (size_t)s * nb;
      const double2* vs = v + (size_t)s * nb;
      double fr = 0.0, fi = 0.0;  // u^T M v, complex, unconjugated
      for (int d = 0; d < 3; ++d) {
        double2 vv[4];
        for (int k = 0; k < 4; ++k) vv[k] = vs[e[4 * d + k]];
        for (int i = 0; i < 4; ++i) {
          const double2 uu = us[e[4 * d + i]];
          double tr = 0.0, ti = 0.0;
          for (int k = 0; k < 4; ++k) {
            tr += m[d][i][k] * vv[k].x;
            ti += m[d][i][k] * vv[k].y;
          }
          fr += uu.x * tr - uu.y * ti;
          fi += uu.x * ti + uu.y * tr;
        }
      }
      g += coef[2 * s] * fr - coef[2 * s + 1] * fi;
    }
    grad[c] = g;
  }
}

// D_A^-1 = 1 / (a + i*omega*b) from the FP64 matrix diagonal (FoldedPreconditioner.cpp:507-520),
// evaluated where it is used (16 bytes per row, shared by the batch) instead of being stored
// as a per-system vector that every application would read.
struct DiagA {
  const double *a, *b, *omega;
};

template <class T>
__device__ inline typename Vec2<T>::type diag_inv(const DiagA& D, int row, int s) {
#pragma clang fp contract(off)
  double qr, qi;
  cdiv(1.0, 0.0, D.a[row], D.omega[s] * D.b[row], qr, qi);
  typename Vec2<T>::type o;
  o.x = T(qr);
  o.y = T(qi);
  return o;
}

// entry (j, c) of the sliced ELL as stored: a and b of a complex-section entry (cplx) or a of a
// real-section one (b = 0); false when absent
__device__ inline bool ell_lookup_ab(const EllView<double>& A, int j, int c, double& va,
                                     double& vb, bool& cplx) {
  const int slice = j >> 6, lane = j & 63;
  const int oc = A.offc[slice], ncs = A.nc[slice];
  for (int q = 0; q < ncs; ++q) {
    const int idx = oc + q * 64 + lane;
    if (A.colc[idx] == c) {
      va = A.ac[idx];
      vb = A.bc[idx];
      cplx = true;
      return true;
    }
  }
  const int orr = A.offr[slice], nrs = A.nr[slice];
  for (int q = 0; q < nrs; ++q) {
    const int idx = orr + q * 64 + lane;
    if (A.colr[idx] == c) {
      va = A.ar[idx];
      vb = 0.0;
      cplx = false;
      return true;
    }
  }
  return false;
}

// D_N^-1 = 1 / (G A G^T)_ii from the FP64 matrix (FoldedPreconditioner.cpp:527-568). The sums run
// in the reference's order (for every edge j of node i ascending: the entries of row j in
// ascending column order, then times G_ij) without fused multiply-adds. In air cells the real part
// of (G A G^T)_ii is the rounding residue of G K G^T = 0, so this order is what fixes its value.
//
// One thread per node: the node's matrix entries are looked up once and its SP systems' sums are
// formed from them (up to 16 systems per pass over the entries, each system's additions in
// exactly the order above), instead of one thread per node and system repeating the lookups.
template <class T, int SP>
__global__ void k_diag_n(EllView<double> A, GView<double> G, const double* omega,
                         typename Vec2<T>::type* out) {
#pragma clang fp contract(off)
  constexpr int CH = SP < 16 ? SP : 16;  // systems per pass over the node's entries
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < G.m;
       i += (long long)gridDim.x * blockDim.x) {
    int ecol[kMaxGRow];
    double eval[kMaxGRow];
    int wg = 0;
    for (int k = 0; k < G.wg && k < kMaxGRow; ++k) {
      const double g = G.gval[(size_t)k * G.m + i];
      if (g == 0.0) continue;  // padding of shorter rows
      ecol[wg] = G.gcol[(size_t)k * G.m + i];
      eval[wg] = g;
      ++wg;
    }
    for (int s0 = 0; s0 < SP; s0 += CH) {
      double w[CH], dr[CH], di[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        w[c] = omega[s0 + c];
        dr[c] = 0.0;
        di[c] = 0.0;
      }
      for (int kj = 0; kj < wg; ++kj) {
        const int j = ecol[kj];
        double sr[CH], si[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) sr[c] = si[c] = 0.0;
        for (int kk = 0; kk < wg; ++kk) {  // ascending edge = ascending column of row j
          double va, vb;
          bool cplx;
          if (ell_lookup_ab(A, j, ecol[kk], va, vb, cplx)) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              const double vi = cplx ? w[c] * vb : 0.0;
              sr[c] += va * eval[kk];
              si[c] += vi * eval[kk];
            }
          }
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          dr[c] += sr[c] * eval[kj];
          di[c] += si[c] * eval[kj];
        }
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        double qr, qi;
        cdiv(1.0, 0.0, dr[c], di[c], qr, qi);
        typename Vec2<T>::type o;
        o.x = T(qr);
        o.y = T(qi);
        out[i * SP + s0 + c] = o;
      }
    }
  }
}

// Output of one (row, system) of the SpMV kernels: y = A x with the partial (y, x) (MODE 0), or
// r = b - A x with the partials ||b||^2, ||r||^2 and the MRS initialisation (MODE 1, 2).
template <class T, int MODE>
__device__ inline void spmv_store(long long t, double yr, double yi,
                                  const typename Vec2<T>::type* __restrict__ xsrc, long long tx,
                                  typename Vec2<T>::type* __restrict__ y,
                                  const typename Vec2<T>::type* __restrict__ b,
                                  typename Vec2<T>::type* __restrict__ ym,
                                  typename Vec2<T>::type* __restrict__ sm, double (&v)[2]) {
  using T2 = typename Vec2<T>::type;
  if (MODE == 0) {
    T2 o;
    o.x = T(yr);
    o.y = T(yi);
    y[t] = o;
    const T2 xt = xsrc[tx];
    // the dot uses the stored (rounded) product, as the reference's ScalCmplx(z, s) does
    v[0] += double(o.x) * double(xt.x) - double(o.y) * double(xt.y);
    v[1] += double(o.x) * double(xt.y) + double(o.y) * double(xt.x);
  } else {
    const T2 bv = b[t];
    T2 o;
    o.x = T(double(bv.x) - yr);
    o.y = T(double(bv.y) - yi);
    y[t] = o;
    v[0] += double(bv.x) * double(bv.x) + double(bv.y) * double(bv.y);
    v[1] += double(o.x) * double(o.x) + double(o.y) * double(o.y);
    if (ym) ym[t] = xsrc[tx];
    if (sm) sm[t] = o;
  }
}

// y = A x (MODE 0, partial (y, x) unconjugated) or r = b - A x (MODE 1, partials ||b||^2 and
// ||r||^2, plus the MRS initialisation ym = x, sm = r of MRS.cpp:91-97). MODE 2 is MODE 1 gated
// by the outer (refinement) status instead of the COCR status.
template <class T, int SP, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(EllView<double> A, const double* omega,
                                                 const typename Vec2<T>::type* __restrict__ x,
                                                 typename Vec2<T>::type* __restrict__ y,
                                                 const typename Vec2<T>::type* __restrict__ b,
                                                 typename Vec2<T>::type* __restrict__ ym,
                                                 typename Vec2<T>::type* __restrict__ sm,
                                                 const SysScalars* __restrict__ sc,
                                                 const int* __restrict__ nrun, double* partials) {
  // The matrix is always applied from its FP64 values with FP64 accumulation, also when the
  // vectors are FP32: rounding K to FP32 breaks the exact discrete-gradient null space
  // (G K = 0) that the folded preconditioner inverts in the near-insulating air, and COCR then
  // stagnates. Rounding happens only when the product is stored. The matrix streams through
  // once per launch, so its loads carry the nontemporal hint (x keeps the caches); eight
  // entries in flight per lane (1 system: 0.214 -> 0.188 ms on a 1.84 M-edge mesh; 12 or 16
  // run slower).
  using T2 = typename Vec2<T>::type;
  if (*nrun == 0) return;
  const long long N = (long long)A.nb * SP;
  const long long ntiles = (N + kBlock - 1) / kBlock;
  double v[2] = {0.0, 0.0};
  TileWalk tw(ntiles);
  const int s = int(threadIdx.x % SP);  // kBlock is a multiple of SP
  const bool run = (MODE == 2 ? sc[s].outer_status : sc[s].status) == ST_RUNNING;
  const double w = omega[s];
  // MODE 0: lanes of a system not running compute and store a dead product too, so that the
  // wave's memory instructions keep every lane active
  for (long long tile = tw.first; (run || MODE == 0) && tile < tw.last; tile += tw.step) {
    const long long t = tile * kBlock + threadIdx.x;
    if (t >= N) continue;
    const int row = int(t / SP);
    const int slice = row >> 6, lane = row & 63;
    double yr = 0, yi = 0;
    {
      const int oc = A.offc[slice], ncs = A.nc[slice];
#pragma unroll 8
      for (int k = 0; k < ncs; ++k) {
        const int idx = oc + k * 64 + lane;
        const int c = __builtin_nontemporal_load(A.colc + idx);
        const double a = __builtin_nontemporal_load(A.ac + idx);
        const double bb = __builtin_nontemporal_load(A.bc + idx) * w;
        const T2 xv = x[(size_t)c * SP + s];
        const double xr = double(xv.x), xi = double(xv.y);
        yr += a * xr - bb * xi;
        yi += a * xi + bb * xr;
      }
    }
    {
      const int orr = A.offr[slice], nrs = A.nr[slice];
#pragma unroll 8
      for (int k = 0; k < nrs; ++k) {
        const int idx = orr + k * 64 + lane;
        const int c = __builtin_nontemporal_load(A.colr + idx);
        const double a = __builtin_nontemporal_load(A.ar + idx);
        const T2 xv = x[(size_t)c * SP + s];
        yr += a * double(xv.x);
        yi += a * double(xv.y);
      }
    }
    spmv_store<T, MODE>(t, yr, yi, x, t, y, b, ym, sm, v);
  }
  block_partials<SP, 2>(v, partials);
}

// Broadcast of lane J of each SP-lane row group to the whole group, for 32-bit values and
// doubles (one 64-bit DPP move): DPP quad_perm (SP 4), DPP row_newbcast within 16-lane rows
// (SP 8, 16), v_readlane (SP 32, 64). Every lane of a group is active at the call (groups are
// whole rows). The DPP forms run with full row and bank masks and write every lane, so their
// `old` operand is never read and is left undefined (no move to initialise it). bound_ctrl is
// set: no source lane is ever out of range or disabled, so it changes no value, and it lets the
// compiler fold a 32-bit broadcast into the add that consumes it (v_add_u32_dpp).
// (Folding the 64-bit broadcasts into v_fmac_f64_dpp as well, by inline asm, cut the VALU count
// of the row-block SpMV by a third but made it slower: 0.80 vs 0.77 ms in FP64, 0.75 vs 0.61 ms
// with FP32 vectors, at 16 systems on a 1.84 M-edge mesh.)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wuninitialized"
template <int SP, int J, class V>
__device__ inline V grp_bcast(V v) {
  static_assert(sizeof(V) == 4 || sizeof(V) == 8, "32- or 64-bit values");
  if constexpr (SP == 4 || SP == 16) {
    constexpr int ctl = SP == 4 ? (J | (J << 2) | (J << 4) | (J << 6)) : 0x150 + J;
    V old;
    return __builtin_amdgcn_update_dpp(old, v, ctl, 0xF, 0xF, true);
  } else if constexpr (SP == 8) {
    V old;
    const V lo = __builtin_amdgcn_update_dpp(old, v, 0x150 + J, 0xF, 0xF, true);
    const V hi = __builtin_amdgcn_update_dpp(old, v, 0x150 + 8 + J, 0xF, 0xF, true);
    return (threadIdx.x & 8) ? hi : lo;
  } else if constexpr (sizeof(V) == 8) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = grp_bcast<SP, J>(int(u & 0xffffffffLL));
    const int hi = grp_bcast<SP, J>(int(u >> 32));
    return __builtin_bit_cast(V, ((long long)hi << 32) | (unsigned int)lo);
  } else if constexpr (SP == 32) {
    const int lo = __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), J);
    const int hi = __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32 + J);
    return __builtin_bit_cast(V, (threadIdx.x & 32) ? hi : lo);
  } else {
    static_assert(SP == 64, "row groups of 4..64 lanes");
    return __builtin_bit_cast(V, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), J));
  }
}
#pragma clang diagnostic pop

template <int J, int N, class F>
__device__ inline void static_for(F&& f) {
  if constexpr (J < N) {
    f(std::integral_constant<int, J>{});
    static_for<J + 1, N>(f);
  }
}

// One chunk of up to SP entries of a row. Lane l of the group holds entry k0 + l: the byte
// offset of its column's SP x values (`off`), a, and b (0 for real entries); the first `nc`
// entries of the chunk are complex. The entries are broadcast one at a time and every lane
// gathers its own system's x through `gather(offset)`. A full chunk with no complex entry
// skips the b terms (a group-uniform branch taken per chunk, so that inside a chunk all SP
// gathers are in flight together); per entry a lane then spends one 32-bit and one 64-bit
// broadcast, the gather and two FMAs, and a complex chunk another broadcast, a multiply and
// two FMAs. Gathers are issued BATCH at a time (a scheduling barrier between batches), which
// bounds the bytes a wave has in flight.
template <class T, int SP, int BATCH, class Gather>
__device__ inline void row_chunk(int cnt, int nc, unsigned off, double a, double bb, double w,
                                 const Gather& gather, double& yr, double& yi) {
  static_assert(SP % BATCH == 0, "whole batches");
  auto step = [&](auto J, auto cplx) {
    const auto xv = gather(grp_bcast<SP, J>(off));
    const double xr = double(xv.x), xm = double(xv.y);
    const double aj = grp_bcast<SP, J>(a);
    yr = __builtin_fma(aj, xr, yr);
    yi = __builtin_fma(aj, xm, yi);
    if constexpr (decltype(cplx)::value) {
      const double bj = grp_bcast<SP, J>(bb) * w;
      yr = __builtin_fma(-bj, xm, yr);
      yi = __builtin_fma(bj, xr, yi);
    }
  };
  using Real = std::false_type;
  using Cplx = std::true_type;
  auto batches = [&](auto kind) {
    static_for<0, SP / BATCH>([&](auto B) {
      if constexpr (B > 0) __builtin_amdgcn_sched_barrier(0);
      static_for<0, BATCH>([&](auto J) {
        step(std::integral_constant<int, B * BATCH + J>{}, kind);
      });
    });
  };
  if (cnt >= SP) {
    if (nc <= 0)
      batches(Real{});
    else
      batches(Cplx{});
  } else {
    static_for<0, SP>([&](auto J) {
      if (J < cnt) step(J, Cplx{});
    });
  }
}

// x of a row kernel as a raw buffer (byte offsets, 32-bit): loads past num_records return 0
// without touching memory, which retired systems use to skip their gathers.
template <class T2>
__device__ inline T2 buffer_gather(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(T2) == 16) {
    return __builtin_bit_cast(T2, __builtin_amdgcn_raw_buffer_load_b128(r, int(off), 0, 0));
  } else {
    static_assert(sizeof(T2) == 8, "float2 or double2");
    return __builtin_bit_cast(T2, __builtin_amdgcn_raw_buffer_load_b64(r, int(off), 0, 0));
  }
}

// the same with a wave-uniform byte offset (the instruction's scalar offset operand)
template <class T2>
__device__ inline T2 buffer_gather_so(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
  if constexpr (sizeof(T2) == 16) {
    return __builtin_bit_cast(
        T2, __builtin_amdgcn_raw_buffer_load_b128(r, int(off), int(soff), 0));
  } else {
    static_assert(sizeof(T2) == 8, "float2 or double2");
    return __builtin_bit_cast(T2, __builtin_amdgcn_raw_buffer_load_b64(r, int(off), int(soff), 0));
  }
}

constexpr int kBufferDword3 = 0x00020000;  // raw buffer descriptor word 3 (CDNA)
// A byte offset past every buffer built by buf_rsrc: raw buffer loads there return 0 and
// touch no memory, which replaces the branch around a load that must not happen.
constexpr unsigned kOob = 0xFFFFFF00u;

__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, long long bytes) {
  const unsigned n = bytes < (long long)kOob ? unsigned(bytes) : kOob;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, int(n),
                                           kBufferDword3);
}
__device__ inline unsigned buf_u32(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, int(off), 0, 0);
}
__device__ inline unsigned buf_u16(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b16(r, int(off), 0, 0);
}
__device__ inline double buf_f64(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(off), 0, 0));
}

// Matrix data of one row for its lane group: the row's bounds and, per lane, entries beg + s
// and beg + SP + s (the first two chunks; b = 0 for real entries and past the end), their
// columns as byte offsets into x (column * SP values).
struct RgRow {
  int beg, end, bbeg, ncx;
  unsigned o0, o1;
  double a0, b0, a1, b1;
};

template <class T, int SP>
__device__ inline void rg_load_row(const RowView& A, int row, int s, RgRow& R) {
  constexpr unsigned kStride = SP * sizeof(typename Vec2<T>::type);
  R.beg = A.rp[row];
  R.end = A.rp[row + 1];
  R.bbeg = A.rb[row];
  R.ncx = A.rb[row + 1] - R.bbeg;
  R.o0 = R.o1 = 0;
  R.a0 = R.b0 = R.a1 = R.b1 = 0.0;
  const int k0 = R.beg + s, k1 = R.beg + SP + s;
  if (k0 < R.end) {
    R.o0 = unsigned(A.col[k0]) * kStride;
    R.a0 = A.a[k0];
    if (s < R.ncx) R.b0 = A.b[R.bbeg + s];
  }
  if (k1 < R.end) {
    R.o1 = unsigned(A.col[k1]) * kStride;
    R.a1 = A.a[k1];
    if (SP + s < R.ncx) R.b1 = A.b[R.bbeg + SP + s];
  }
}

// Row-group SpMV over the row pack: SP lanes per row (one per system), the row's entries
// loaded SP at a time with one coalesced load per lane and broadcast inside the group, so the
// matrix costs one load per entry per group instead of one per entry per lane. The first two
// chunks of a row (32 entries at SP = 16: a whole interior row of a hexahedral mesh) are
// loaded together, and the next tile's are loaded while the current tile's gathers are in
// flight, so that the dependent row-pointer -> entry -> gather chains overlap. The gathers use
// 32-bit byte offsets from x (the host keeps x under 4 GiB for this kernel) and are issued four
// at a time: with all 16 of a chunk in flight at once the FP64 kernel ran at half the speed
// (0.91 -> 1.84 ms, 16 systems on 1.84 M rows). Same modes, outputs and partials as k_spmv;
// the diagonal is applied first.
template <class T, int SP, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_rg(RowView A, const double* omega,
                                                    const typename Vec2<T>::type* __restrict__ x,
                                                    typename Vec2<T>::type* __restrict__ y,
                                                    const typename Vec2<T>::type* __restrict__ b,
                                                    typename Vec2<T>::type* __restrict__ ym,
                                                    typename Vec2<T>::type* __restrict__ sm,
                                                    const SysScalars* __restrict__ sc,
                                                    const int* __restrict__ nrun,
                                                    double* partials) {
  static_assert(SP >= 4 && kBlock % SP == 0, "row groups of 4..64 lanes");
  using T2 = typename Vec2<T>::type;
  if (*nrun == 0) return;
  const long long N = (long long)A.nb * SP;
  const long long ntiles = (N + kBlock - 1) / kBlock;
  double v[2] = {0.0, 0.0};
  TileWalk tw(ntiles);
  const int s = int(threadIdx.x % SP);  // kBlock is a multiple of SP
  const bool run = (MODE == 2 ? sc[s].outer_status : sc[s].status) == ST_RUNNING;
  const double w = omega[s];
  const unsigned char* xb = reinterpret_cast<const unsigned char*>(x);
  const unsigned s_off = unsigned(s * sizeof(T2));
  auto gather = [&](unsigned o) { return *reinterpret_cast<const T2*>(xb + (o + s_off)); };
  // N is a multiple of SP: whole groups are in or out of range together
  auto row_of = [&](long long tile) { return tile * kBlock + threadIdx.x < N
                                          ? int((tile * kBlock + threadIdx.x) / SP) : -1; };
  RgRow cur, nxt;
  long long tile = tw.first;
  if (tile < tw.last) {
    const int r0 = row_of(tile);
    if (r0 >= 0) rg_load_row<T, SP>(A, r0, s, cur);
  }
  for (; tile < tw.last; tile += tw.step) {
    const int row = row_of(tile);
    const long long ntile = tile + tw.step;
    if (ntile < tw.last) {
      const int nrow = row_of(ntile);
      if (nrow >= 0) rg_load_row<T, SP>(A, nrow, s, nxt);
    }
    if (row >= 0) {
      const long long t = (long long)row * SP + s;
      const T2 xs = x[t];
      const double dA = A.da[row], dB = A.db[row] * w;
      double yr = dA * double(xs.x) - dB * double(xs.y);
      double yi = dA * double(xs.y) + dB * double(xs.x);
      const int len = cur.end - cur.beg;
      row_chunk<T, SP, 4>(len, cur.ncx, cur.o0, cur.a0, cur.b0, w, gather, yr, yi);
      if (len > SP)
        row_chunk<T, SP, 4>(len - SP, cur.ncx - SP, cur.o1, cur.a1, cur.b1, w, gather, yr,
                             yi);
      for (int k0 = cur.beg + 2 * SP; k0 < cur.end; k0 += SP) {  // rows longer than 2 chunks
        const int l = k0 - cur.beg + s;
        unsigned o = 0;
        double a = 0.0, bb = 0.0;
        if (k0 + s < cur.end) {
          o = unsigned(A.col[k0 + s]) * unsigned(SP * sizeof(T2));
          a = A.a[k0 + s];
          if (l < cur.ncx) bb = A.b[cur.bbeg + l];
        }
        row_chunk<T, SP, 4>(cur.end - k0, cur.ncx - (k0 - cur.beg), o, a, bb, w, gather,
                             yr, yi);
      }
      if (run || MODE == 0) spmv_store<T, MODE>(t, yr, yi, x, t, y, b, ym, sm, v);
    }
    cur = nxt;
  }
  block_partials<SP, 2>(v, partials);
}

// Halo-staged SpMV over 16-row blocks (batch widths 4..16). A workgroup copies the x values
// of a block's distinct columns (its halo, every system) into LDS with coalesced loads, then
// each row group (SP lanes, one per system) walks its row as k_spmv_rg does, reading x from LDS
// through the entries' 16-bit halo positions. x then crosses the vector memory path once per
// block column instead of once per non-zero. 256/SP row groups cover 16/SP blocks at a time,
// each with its own LDS region of hmax * SP values (hmax <= kHaloStage). The walk is software
// pipelined: while a tile computes from LDS, the next tile's x values and row data and the
// tile after's halo indices are in flight in registers. Same modes, outputs and partials as
// k_spmv; the diagonal is applied first.
template <class T, int SP, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_blk(BlkView A, int hmax, const double* omega,
                                                     const typename Vec2<T>::type* __restrict__ x,
                                                     typename Vec2<T>::type* __restrict__ y,
                                                     const typename Vec2<T>::type* __restrict__ b,
                                                     typename Vec2<T>::type* __restrict__ ym,
                                                     typename Vec2<T>::type* __restrict__ sm,
                                                     const SysScalars* __restrict__ sc,
                                                     const int* __restrict__ nrun,
                                                     double* partials) {
  static_assert(SP >= 4 && SP <= 16 && kBlock % (SP * kRowBlock) == 0, "4..16 systems");
  using T2 = typename Vec2<T>::type;
  extern __shared__ __align__(16) unsigned char blk_lds[];
  if (*nrun == 0) return;
  constexpr int BW = kBlock / (SP * kRowBlock);  // blocks per pass
  constexpr int BT = SP * kRowBlock;             // threads per block
  constexpr int NI = kHaloStage / kRowBlock;     // staging slots per thread
  const int s = int(threadIdx.x % SP);
  const int sub = int(threadIdx.x / BT);
  const int rl = int(threadIdx.x % BT) / SP;     // row within the block
  T2* xs = reinterpret_cast<T2*>(blk_lds) + (size_t)sub * hmax * SP;
  const unsigned char* xs_s = reinterpret_cast<const unsigned char*>(xs + s);
  auto gather = [&](unsigned o) { return *reinterpret_cast<const T2*>(xs_s + o); };
  constexpr unsigned kStride = SP * sizeof(T2);
  const bool run = (MODE == 2 ? sc[s].outer_status : sc[s].status) == ST_RUNNING;
  const double w = omega[s];
  double v[2] = {0.0, 0.0};
  const long long ntiles = ((long long)A.nblk + BW - 1) / BW;
  TileWalk tw(ntiles);

  // Every load of the staging and the row data is a raw buffer load with a 32-bit offset;
  // loads that must not happen (past the halo, past the row, retired systems) use offset kOob
  // and return 0. The row pack's entry arrays are addressed from the tile's first entry.
  const __amdgpu_buffer_rsrc_t hr = buf_rsrc(A.halo, A.nhalo * 4);
  const __amdgpu_buffer_rsrc_t xr = buf_rsrc(x, (long long)A.nb * SP * (long long)sizeof(T2));
  const __amdgpu_buffer_rsrc_t rpr = buf_rsrc(A.rp, ((long long)A.nb + 1) * 4);
  const __amdgpu_buffer_rsrc_t rbr = buf_rsrc(A.rb, ((long long)A.nb + 1) * 4);
  const __amdgpu_buffer_rsrc_t dlr = buf_rsrc(A.ldiag, (long long)A.nb * 2);
  const __amdgpu_buffer_rsrc_t dar = buf_rsrc(A.da, (long long)A.nb * 8);
  const __amdgpu_buffer_rsrc_t dbr = buf_rsrc(A.db, (long long)A.nb * 8);
  const unsigned s_off = unsigned(s * sizeof(T2));

  // halo slot q of this thread is column rl + 16 q of its block (cm: the slots in the halo)
  unsigned ci[NI];
  unsigned cm = 0;
  auto load_idx = [&](long long tile) {
    const long long blk = tile * BW + sub;
    int h0 = 0, hn = 0;
    if (blk < A.nblk) {
      h0 = A.hoff[blk];
      hn = A.hoff[blk + 1] - h0;
    }
    cm = 0;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const int i = rl + q * kRowBlock;
      ci[q] = buf_u32(hr, i < hn ? unsigned(h0 + i) * 4u : kOob);
      cm |= (i < hn ? 1u : 0u) << q;
    }
  };
  T2 xv[NI];
  unsigned xmask = 0;  // slots of the staged tile that hold a halo column
  auto load_x = [&]() {
    xmask = cm;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const bool take = run && (cm & (1u << q));
      xv[q] = buffer_gather<T2>(xr, take ? ci[q] * kStride + s_off : kOob);
    }
  };
  struct Row {
    int beg, end, bbeg, ncx, dl;
    unsigned o0, o1;
    double a0, b0, a1, b1, dA, dB;
  };
  auto load_row = [&](long long tile, Row& R) {
    const long long r0 = tile * BW * kRowBlock;  // the tile's first row (uniform, < nb)
    const long long rowl = r0 + sub * kRowBlock + rl;
    const unsigned row = unsigned(rowl);
    const bool ok = rowl < A.nb;
    const int e0 = A.rp[r0], f0 = A.rb[r0];
    const __amdgpu_buffer_rsrc_t ar = buf_rsrc(A.a + e0, (A.nnz - e0) * 8);
    const __amdgpu_buffer_rsrc_t lr = buf_rsrc(A.lcol + e0, (A.nnz - e0) * 2);
    const __amdgpu_buffer_rsrc_t br = buf_rsrc(A.b + f0, (A.nnzb - f0) * 8);
    R.beg = int(buf_u32(rpr, ok ? row * 4u : kOob));
    R.end = int(buf_u32(rpr, ok ? row * 4u + 4u : kOob));
    R.bbeg = int(buf_u32(rbr, ok ? row * 4u : kOob));
    R.ncx = int(buf_u32(rbr, ok ? row * 4u + 4u : kOob)) - R.bbeg;
    R.dl = int(buf_u16(dlr, ok ? row * 2u : kOob));
    R.dA = buf_f64(dar, ok ? row * 8u : kOob);
    R.dB = buf_f64(dbr, ok ? row * 8u : kOob) * w;
    const int len = R.end - R.beg;
    const unsigned k0 = unsigned(R.beg - e0 + s), k1 = k0 + SP;
    const unsigned q0 = unsigned(R.bbeg - f0 + s), q1 = q0 + SP;
    R.o0 = buf_u16(lr, s < len ? k0 * 2u : kOob) * kStride;
    R.a0 = buf_f64(ar, s < len ? k0 * 8u : kOob);
    R.b0 = buf_f64(br, s < R.ncx ? q0 * 8u : kOob);
    R.o1 = buf_u16(lr, SP + s < len ? k1 * 2u : kOob) * kStride;
    R.a1 = buf_f64(ar, SP + s < len ? k1 * 8u : kOob);
    R.b1 = buf_f64(br, SP + s < R.ncx ? q1 * 8u : kOob);
  };

  Row cur, nxt;
  long long tile = tw.first;
  if (tile < tw.last) {
    load_idx(tile);
    load_x();
    load_row(tile, cur);
    if (tile + tw.step < tw.last) load_idx(tile + tw.step);
  }
  for (; tile < tw.last; tile += tw.step) {
#pragma unroll
    for (int q = 0; q < NI; ++q)
      if (xmask & (1u << q)) xs[(rl + q * kRowBlock) * SP + s] = xv[q];
    __syncthreads();
    const long long nt = tile + tw.step;
    if (nt < tw.last) {  // ci holds tile nt's halo
      load_x();
      load_row(nt, nxt);
      if (nt + tw.step < tw.last) load_idx(nt + tw.step);
    }
    const long long rowl = (tile * BW + sub) * kRowBlock + rl;
    if (rowl < A.nb) {
      const long long t = rowl * SP + s;
      const T2 xt = xs[cur.dl * SP + s];
      double yr = cur.dA * double(xt.x) - cur.dB * double(xt.y);
      double yi = cur.dA * double(xt.y) + cur.dB * double(xt.x);
      const int len = cur.end - cur.beg;
      row_chunk<T, SP, SP>(len, cur.ncx, cur.o0, cur.a0, cur.b0, w, gather, yr, yi);
      if (len > SP)
        row_chunk<T, SP, SP>(len - SP, cur.ncx - SP, cur.o1, cur.a1, cur.b1, w, gather, yr,
                             yi);
      for (int k0 = cur.beg + 2 * SP; k0 < cur.end; k0 += SP) {  // rows longer than 2 chunks
        const int l = k0 - cur.beg + s;
        unsigned o = 0;
        double a = 0.0, bb = 0.0;
        if (k0 + s < cur.end) {
          o = unsigned(A.lcol[k0 + s]) * kStride;
          a = A.a[k0 + s];
          if (l < cur.ncx) bb = A.b[cur.bbeg + l];
        }
        row_chunk<T, SP, SP>(cur.end - k0, cur.ncx - (k0 - cur.beg), o, a, bb, w, gather,
                             yr, yi);
      }
      if (run || MODE == 0) spmv_store<T, MODE>(t, yr, yi, xs, cur.dl * SP + s, y, b, ym, sm, v);
    }
    cur = nxt;
    __syncthreads();  // the next tile's staging overwrites xs
  }
  block_partials<SP, 2>(v, partials);
}

// Structured-mesh stencil SpMV (batch widths 4..16) over the grid order of a box mesh (rows:
// the x-, y- and z-edge boxes, each in (k, j, i) order; host_prep.h). Every row couples with a
// fixed 33-point stencil whose positions follow from the row's (direction, i, j, k), so no
// column indices are stored and no halo is staged: the matrix is 42 FP64 coefficients per row.
// A wave owns RPW = 64/SP consecutive rows in i (SP lanes per row, one per system) of one
// direction and one j, and sweeps a chunk of k. The stencil's points are grouped in columns
// (direction, di, dj) spanning two or three k levels; each value x of column c at level L is
// gathered once (one coalesced SP-value segment per row) and pushed into the rows L - dk it
// couples with, so the sweep gathers 13 values per row and level (21 for z-edges) instead of 33.
// Rows L-1..L+2 keep accumulators and their coefficients (spread over the row's SP lanes,
// broadcast by DPP at use) in a ring of four; the next level's gathers and the coefficients two
// rows ahead are issued before the current level's arithmetic. Neighbour positions are clamped
// into their box: a row with a missing neighbour (a boundary row) has a zero coefficient there.
// Same modes, outputs and partials as k_spmv.
// Products with a coefficient broadcast from lane J of each 16-lane row, the broadcast folded
// into the FMA (v_fmac_f64_dpp, row_newbcast). Written as inline asm because the compiler,
// given separate DPP moves, schedules all of a level's broadcasts ahead of their FMAs and then
// runs out of registers. The compiler does not count wait states for inline asm, so each block
// starts with the two that a DPP read of a VGPR written by a VALU instruction needs (s_nop 1);
// inside a block no instruction writes a DPP source.
//   ar += c xr, ai += c xi
template <int J>
__device__ inline void dfma2_bc(double& ar, double& ai, double c, double xr, double xi) {
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, %2, %3 row_newbcast:%5 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f64_dpp %1, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(ar), "+v"(ai)
      : "v"(c), "v"(xr), "v"(xi), "i"(J));
}
//   (ar, ai) += (ca + i w cb)(xr + i xi), with nwxi = -w xi and wxr = w xr given
template <int Ja, int Jb>
__device__ inline void dfma4_bc(double& ar, double& ai, double ca, double cb, double xr, double xi,
                                double nwxi, double wxr) {
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f64_dpp %1, %2, %5 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f64_dpp %0, %3, %6 row_newbcast:%9 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f64_dpp %1, %3, %7 row_newbcast:%9 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(ar), "+v"(ai)
      : "v"(ca), "v"(cb), "v"(xr), "v"(xi), "v"(nwxi), "v"(wxr), "i"(Ja), "i"(Jb));
}

// One component per lane (the 8-system stencil kernel, whose 16-lane rows hold 8 systems x
// (re, im)): acc += c u, and acc += ca u + cb v, with the same broadcasts and FMA order as the
// component of dfma2_bc / dfma4_bc that the lane stands for.
template <int J>
__device__ inline void dfma1_bc(double& acc, double c, double u) {
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(acc)
      : "v"(c), "v"(u), "i"(J));
}
template <int Ja, int Jb>
__device__ inline void dfma2x_bc(double& acc, double ca, double cb, double u, double v) {
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, %1, %3 row_newbcast:%5 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%6 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(acc)
      : "v"(ca), "v"(cb), "v"(u), "v"(v), "i"(Ja), "i"(Jb));
}

struct StcView {
  const double* coef;  // nb x kStcCoef, row-major
  int nx, ny, nz;      // nodes per axis
  int kc;              // rows of one sweep (k-chunk)
  int nic;             // i-chunks of RPW rows
  long long ntasks;    // 3 * nkc * ny * nic wave tasks (k_spmv_stc), 3 * nkc * njc * nic
                       // workgroup tasks (k_spmv_stl)
  int njc;             // j-chunks of kStlNW rows (k_spmv_stl)
  int nbw;             // k_spmv_stl: tile columns per band of the task order
  int nb;
};

// Output of one (row, system) of the stencil kernel, as spmv_store, through raw buffers: when
// `ok` is false the offset is kOob, so the stores are dropped, the loads return 0 and the
// partial sums receive nothing.
template <class T, int MODE>
__device__ inline void stc_store(bool ok, unsigned off, double yr, double yi,
                                 __amdgpu_buffer_rsrc_t xr, __amdgpu_buffer_rsrc_t y_r,
                                 __amdgpu_buffer_rsrc_t b_r, __amdgpu_buffer_rsrc_t ym_r,
                                 __amdgpu_buffer_rsrc_t sm_r, double (&v)[2]) {
  using T2 = typename Vec2<T>::type;
  const unsigned o = ok ? off : kOob;
  auto store = [&](__amdgpu_buffer_rsrc_t r, T2 val) {
    if constexpr (sizeof(T2) == 16)
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, val), r, int(o), 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, val), r, int(o), 0, 0);
  };
  if (MODE == 0) {
    T2 out;
    out.x = T(yr);
    out.y = T(yi);
    store(y_r, out);
    const T2 xt = buffer_gather<T2>(xr, o);  // 0 when the row is not stored
    // the dot uses the stored (rounded) product, as the reference's ScalCmplx(z, s) does
    v[0] += double(out.x) * double(xt.x) - double(out.y) * double(xt.y);
    v[1] += double(out.x) * double(xt.y) + double(out.y) * double(xt.x);
  } else {
    const T2 bv = buffer_gather<T2>(b_r, o);
    T2 out;
    out.x = T(double(bv.x) - yr);
    out.y = T(double(bv.y) - yi);
    store(y_r, out);
    v[0] += double(bv.x) * double(bv.x) + double(bv.y) * double(bv.y);
    const double rr = double(out.x) * double(out.x) + double(out.y) * double(out.y);
    v[1] += ok ? rr : 0.0;
    const T2 xt = buffer_gather<T2>(xr, o);
    store(ym_r, xt);
    store(sm_r, out);
  }
}

// stc_store with the row's own x given (the LDS-staged kernel has it from the previous level)
// instead of loaded: the same outputs and partial sums, bit for bit
template <class T, int MODE>
__device__ inline void stl_store(bool ok, unsigned off, double yr, double yi,
                                 typename Vec2<T>::type xself, __amdgpu_buffer_rsrc_t y_r,
                                 __amdgpu_buffer_rsrc_t b_r, __amdgpu_buffer_rsrc_t ym_r,
                                 __amdgpu_buffer_rsrc_t sm_r, double (&v)[2]) {
  using T2 = typename Vec2<T>::type;
  const unsigned o = ok ? off : kOob;
  auto store = [&](__amdgpu_buffer_rsrc_t r, T2 val) {
    if constexpr (sizeof(T2) == 16)
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, val), r, int(o), 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, val), r, int(o), 0, 0);
  };
  T2 xt;
  xt.x = ok ? xself.x : T(0);
  xt.y = ok ? xself.y : T(0);
  if (MODE == 0) {
    T2 out;
    out.x = T(yr);
    out.y = T(yi);
    store(y_r, out);
    v[0] += double(out.x) * double(xt.x) - double(out.y) * double(xt.y);
    v[1] += double(out.x) * double(xt.y) + double(out.y) * double(xt.x);
  } else {
    const T2 bv = buffer_gather<T2>(b_r, o);
    T2 out;
    out.x = T(double(bv.x) - yr);
    out.y = T(double(bv.y) - yi);
    store(y_r, out);
    v[0] += double(bv.x) * double(bv.x) + double(bv.y) * double(bv.y);
    const double rr = double(out.x) * double(out.x) + double(out.y) * double(out.y);
    v[1] += ok ? rr : 0.0;
    store(ym_r, xt);
    store(sm_r, out);
  }
}

template <class T, int SP, int MODE, int D>
__device__ inline void stc_sweep(const StcView& A, int ic, int j, int kc, int s, int g, double w,
                                 bool run, __amdgpu_buffer_rsrc_t xr, __amdgpu_buffer_rsrc_t cr,
                                 __amdgpu_buffer_rsrc_t yr_, __amdgpu_buffer_rsrc_t br_,
                                 __amdgpu_buffer_rsrc_t ymr, __amdgpu_buffer_rsrc_t smr,
                                 double (&v)[2]) {
  using T2 = typename Vec2<T>::type;
  constexpr int RPW = 64 / SP;
  constexpr int NC = kStcCols[D];
  constexpr int NCR = (kStcCoef + SP - 1) / SP;  // coefficient registers per lane and row
  constexpr int LO = D == 2 ? 0 : -1, HI = 1;    // level range of the stencil (dk)
  constexpr unsigned kRow = SP * sizeof(T2);      // bytes of one row of x (all systems)
  const GridDims G{A.nx, A.ny, A.nz};
  const int nxd = G.nxd(D), nyd = G.nyd(D), nzd = G.nzd(D);
  const int i = ic * RPW + g;
  const int k0 = kc * A.kc, k1 = min(k0 + A.kc, nzd);
  if (j >= nyd || k0 >= nzd || ic * RPW >= nxd) return;  // the whole task is empty
  const bool row_ok = i < nxd;
  const int ir = min(i, nxd - 1);
  // per column: the lane's byte offset in x (its i clamped into the column's box), the
  // uniform row stride of one level, and the uniform part of the row index
  unsigned vo[NC];
  long long ub[NC], us[NC];
  int nzc[NC];
  static_for<0, NC>([&](auto C) {
    constexpr StencilCol c = kStc[D][C];
    const int nxc = G.nxd(c.dt), nyc = G.nyd(c.dt);
    const int ic_ = min(max(i + c.di, 0), nxc - 1);
    const int jc = min(max(j + c.dj, 0), nyc - 1);
    vo[C] = unsigned(ic_) * kRow + unsigned(s) * unsigned(sizeof(T2));
    ub[C] = G.base(c.dt) + (long long)jc * nxc;
    us[C] = (long long)nyc * nxc;
    nzc[C] = G.nzd(c.dt);
  });
  // the value of column C at level L (the level clamped into the column's box)
  auto gather1 = [&](auto Cc, int L) {
    constexpr int C = decltype(Cc)::value;
    const int Lc = min(max(L, 0), nzc[C] - 1);
    const unsigned so = unsigned((ub[C] + (long long)Lc * us[C]) * kRow);
    return buffer_gather_so<T2>(xr, vo[C], so);
  };
  // coefficients of row k: lane s of the row holds slots s, s + SP, ... (0 past the row)
  auto load_coef = [&](double (&dst)[NCR], int k) {
    const int kk = min(max(k, 0), nzd - 1);
    const long long r = G.base(D) + ((long long)kk * nyd + j) * nxd + ir;
    const unsigned rb = unsigned(r) * unsigned(kStcCoef * 8);
    static_for<0, NCR>([&](auto Q) {
      const int slot = Q * SP + s;
      dst[Q] = buf_f64(cr, slot < kStcCoef ? rb + unsigned(slot) * 8u : kOob);
    });
  };
  // Columns are processed in batches of NBC; the values of the next NBUF - 1 batches (across
  // level boundaries) are in flight while a batch computes, and scheduling barriers keep the
  // compiler from hoisting more loads than that (which would cost occupancy). Batch b of level
  // L0 + n is batch g = n * NB + b of the sweep and lives in buffer g mod NBUF.
  constexpr int NBC = 4;
  constexpr int NBUF = 4;
  constexpr int NB = (NC + NBC - 1) / NBC;  // batches per level
  T2 vb[NBUF][NBC];
  double cf[4][NCR];
  double ar[4], ai[4];
  const int L0 = k0 + LO, L1 = k1 - 1 + HI;  // levels the rows k0..k1-1 need
  auto load_batch = [&](T2 (&dst)[NBC], auto Bc, int L) {
    constexpr int B = decltype(Bc)::value;
    static_for<0, NBC>([&](auto Jc) {
      constexpr int C = B * NBC + decltype(Jc)::value;
      if constexpr (C < NC) dst[decltype(Jc)::value] = gather1(std::integral_constant<int, C>{}, L);
    });
  };
  // Rings indexed relative to the first level L0: row L0 + m sits in slot m mod 4, so every
  // index below is a compile-time constant (NBUF divides 4, the level unroll).
  static_for<0, NBUF - 1>([&](auto Gc) {
    constexpr int G = decltype(Gc)::value;
    if (L0 + G / NB <= L1) load_batch(vb[G % NBUF], std::integral_constant<int, G % NB>{}, L0 + G / NB);
  });
  load_coef(cf[3], L0 - 1);
  load_coef(cf[0], L0);
  load_coef(cf[1], L0 + 1);
  ar[3] = ai[3] = ar[0] = ai[0] = 0.0;
  // level L = L0 + n with P = n mod 4
  auto level = [&](auto Pc, int L) {
    constexpr int P = decltype(Pc)::value;
    load_coef(cf[(P + 2) % 4], L + 2);   // row L+2, first used at level L+1
    ar[(P + 1) % 4] = ai[(P + 1) % 4] = 0.0;  // row L+1, first contribution at level L
    static_for<0, NB>([&](auto Bc) {
      constexpr int B = decltype(Bc)::value;
      constexpr int g = P * NB + B;                 // mod NBUF, the sweep's batch index
      constexpr int cur = g % NBUF;
      constexpr int ahead = B + NBUF - 1;           // batch to issue, counted from this level
      __builtin_amdgcn_sched_barrier(0);
      if (L + ahead / NB <= L1)
        load_batch(vb[(g + NBUF - 1) % NBUF], std::integral_constant<int, ahead % NB>{},
                   L + ahead / NB);
      __builtin_amdgcn_sched_barrier(0);
      static_for<0, NBC>([&](auto Jc) {
        constexpr int J = decltype(Jc)::value;
        constexpr int C = B * NBC + J;
        if constexpr (C < NC) {
          constexpr StencilCol c = kStc[D][C];
          const double xr_ = double(vb[cur][J].x), xi_ = double(vb[cur][J].y);
          // a parallel column's coefficients are complex: the omega-scaled parts once per value
          const double nwxi = c.dt == D ? -w * xi_ : 0.0, wxr = c.dt == D ? w * xr_ : 0.0;
          static_for<c.lo, c.hi + 1>([&](auto DK) {
            constexpr int dk = decltype(DK)::value;
            constexpr int R = ((P - dk) % 4 + 4) % 4;  // ring slot of row L - dk
            constexpr int o = stc_slot(D, C, dk);
            if constexpr (c.dt == D)
              dfma4_bc<o % SP, (o + 1) % SP>(ar[R], ai[R], cf[R][o / SP], cf[R][(o + 1) / SP],
                                             xr_, xi_, nwxi, wxr);
            else
              dfma2_bc<o % SP>(ar[R], ai[R], cf[R][o / SP], xr_, xi_);
          });
        }
      });
    });
    // row L - HI has all its levels: store it (without a branch, so that the compiler does
    // not sink the row's last products into a conditional block away from their broadcasts)
    const int k = L - HI;
    constexpr int R = ((P - HI) % 4 + 4) % 4;
    const bool ok = k >= k0 && k < k1 && row_ok && (run || MODE == 0);
    const long long t = (G.base(D) + ((long long)min(max(k, 0), nzd - 1) * nyd + j) * nxd + ir) *
                            SP + s;
    stc_store<T, MODE>(ok, unsigned(t) * unsigned(sizeof(T2)), ar[R], ai[R], xr, yr_, br_, ymr,
                       smr, v);
  };
  for (int L = L0; L <= L1; L += 4) {
    level(std::integral_constant<int, 0>{}, L);
    if (L + 1 > L1) break;
    level(std::integral_constant<int, 1>{}, L + 1);
    if (L + 2 > L1) break;
    level(std::integral_constant<int, 2>{}, L + 2);
    if (L + 3 > L1) break;
    level(std::integral_constant<int, 3>{}, L + 3);
  }
}

template <class T, int SP, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_stc(StcView A, const double* omega,
                                                     const typename Vec2<T>::type* __restrict__ x,
                                                     typename Vec2<T>::type* __restrict__ y,
                                                     const typename Vec2<T>::type* __restrict__ b,
                                                     typename Vec2<T>::type* __restrict__ ym,
                                                     typename Vec2<T>::type* __restrict__ sm,
                                                     const SysScalars* __restrict__ sc,
                                                     const int* __restrict__ nrun,
                                                     double* partials) {
  static_assert(SP == 16, "16-lane rows (the DPP row broadcast)");
  using T2 = typename Vec2<T>::type;
  if (*nrun == 0) return;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int s = lane % SP, g = lane / SP;
  const bool run = (MODE == 2 ? sc[s].outer_status : sc[s].status) == ST_RUNNING;
  const double w = omega[s];
  double v[2] = {0.0, 0.0};
  const long long vbytes = (long long)A.nb * SP * (long long)sizeof(T2);
  const __amdgpu_buffer_rsrc_t xr = buf_rsrc(x, vbytes);
  const __amdgpu_buffer_rsrc_t cr = buf_rsrc(A.coef, (long long)A.nb * kStcCoef * 8);
  // output buffers; an absent one (null: b, ym, sm in MODE 0, ym and sm optional in MODE 1)
  // gets zero records, so that its stores are dropped and its loads return 0
  const __amdgpu_buffer_rsrc_t y_r = buf_rsrc(y, vbytes);
  const __amdgpu_buffer_rsrc_t b_r = buf_rsrc(b, b ? vbytes : 0);
  const __amdgpu_buffer_rsrc_t ym_r = buf_rsrc(ym, ym ? vbytes : 0);
  const __amdgpu_buffer_rsrc_t sm_r = buf_rsrc(sm, sm ? vbytes : 0);
  const long long ntiles = (A.ntasks + 3) / 4;
  TileWalk tw(ntiles);
  for (long long tile = tw.first; tile < tw.last; tile += tw.step) {
    const long long task = tile * 4 + wv;
    if (task >= A.ntasks) continue;
    const int d = int(task % 3);
    long long q = task / 3;
    const int ic = int(q % A.nic);
    q /= A.nic;
    const int j = int(q % A.ny);
    const int kc = int(q / A.ny);
    if (d == 0)
      stc_sweep<T, SP, MODE, 0>(A, ic, j, kc, s, g, w, run, xr, cr, y_r, b_r, ym_r, sm_r, v);
    else if (d == 1)
      stc_sweep<T, SP, MODE, 1>(A, ic, j, kc, s, g, w, run, xr, cr, y_r, b_r, ym_r, sm_r, v);
    else
      stc_sweep<T, SP, MODE, 2>(A, ic, j, kc, s, g, w, run, xr, cr, y_r, b_r, ym_r, sm_r, v);
  }
  block_partials<SP, 2>(v, partials);
}

// LDS-staged stencil SpMV (layout 4, 16 systems). A workgroup of kStlNW waves owns a tile of
// rows of one direction: RPW = 4 consecutive rows in i (wave-wide, as in k_spmv_stc) times
// kStlNW consecutive rows in j (one per wave), and sweeps a chunk of k. At each level L the
// tile's x neighbourhood at L (per column direction dt, the rectangle that the tile's columns
// reach) is copied into LDS once, by all 256 threads with coalesced 16-byte loads, and every
// wave then reads its 13 (21) column values from LDS and pushes them into rows L-1..L+1 as
// k_spmv_stc does. Each x value crosses L2 -> CU once per tile and level instead of once per
// column and row group. Two plane buffers: plane L+1 is loaded into registers while level L
// computes and written to the other buffer afterwards, with one workgroup barrier per level.
constexpr int kStlNW = kBlock / 64;  // waves per workgroup = rows of a tile in j
constexpr int kWholeColumn = 1 << 20;  // a stencil sweep length past any mesh's nz
constexpr int kStlTI = 4;            // rows of a tile in i (64 / 16 lanes per row)
constexpr int stl_imin(int a, int b) { return a < b ? a : b; }
constexpr int stl_imax(int a, int b) { return a > b ? a : b; }
// range of di (which 0, 1) or dj (2, 3) over the columns of a type-D row that read dt
constexpr int stl_rng(int D, int dt, int which) {
  int r = which % 2 == 0 ? 99 : -99;
  for (int c = 0; c < kStcCols[D]; ++c)
    if (kStc[D][c].dt == dt) {
      const int v = which < 2 ? kStc[D][c].di : kStc[D][c].dj;
      r = which % 2 == 0 ? stl_imin(r, v) : stl_imax(r, v);
    }
  return r;
}
constexpr int stl_wid(int D, int dt) { return kStlTI + stl_rng(D, dt, 1) - stl_rng(D, dt, 0); }
constexpr int stl_hgt(int D, int dt) { return kStlNW + stl_rng(D, dt, 3) - stl_rng(D, dt, 2); }
constexpr int stl_base(int D, int dt) {
  int b = 0;
  for (int q = 0; q < dt; ++q) b += stl_wid(D, q) * stl_hgt(D, q);
  return b;
}
constexpr int stl_nval(int D) { return stl_base(D, 3); }
// the column of a type-D row's own edge (dt = D, di = dj = 0)
constexpr int stl_self(int D) {
  for (int c = 0; c < kStcCols[D]; ++c)
    if (kStc[D][c].dt == D && kStc[D][c].di == 0 && kStc[D][c].dj == 0) return c;
  return -1;
}
static_assert(stl_self(0) >= 0 && stl_self(1) >= 0 && stl_self(2) >= 0, "own column");
constexpr int kStlMaxVal = stl_imax(stl_nval(0), stl_imax(stl_nval(1), stl_nval(2)));
static_assert(stl_nval(0) == 79 && stl_nval(1) == 79 && stl_nval(2) == 96,
              "plane sizes of a 4 x 4 tile");

// stl_store for one component (comp 0: re, 1: im) of the (row, system) element at byte offset
// `off`, as the 8-system stencil kernel's lanes hold them: the same stored values; each lane
// adds its component's terms of the partial sums (the pair of lanes adds up to stl_store's).
template <class T, int MODE>
__device__ inline void stl_store_split(bool ok, unsigned off, int comp, double acc,
                                       typename Vec2<T>::type xself, __amdgpu_buffer_rsrc_t y_r,
                                       __amdgpu_buffer_rsrc_t b_r, __amdgpu_buffer_rsrc_t ym_r,
                                       __amdgpu_buffer_rsrc_t sm_r, double (&v)[2]) {
  const unsigned o = ok ? off + unsigned(comp) * unsigned(sizeof(T)) : kOob;
  auto store = [&](__amdgpu_buffer_rsrc_t r, T val) {
    if constexpr (sizeof(T) == 8)
      __builtin_amdgcn_raw_buffer_store_b64(
          __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, val), r, int(o), 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, val), r, int(o), 0, 0);
  };
  const T xc = ok ? (comp ? xself.y : xself.x) : T(0);  // this component of the row's own x
  const T xo = ok ? (comp ? xself.x : xself.y) : T(0);  // the other one
  if (MODE == 0) {
    const T out = T(acc);
    store(y_r, out);
    // (z, x) = (zr xr - zi xi) + i (zr xi + zi xr): comp 0 holds zr, comp 1 holds zi
    v[0] += comp ? -double(out) * double(xc) : double(out) * double(xc);
    v[1] += double(out) * double(xo);
  } else {
    T bc;
    if constexpr (sizeof(T) == 8)
      bc = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(b_r, int(o), 0, 0));
    else
      bc = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(b_r, int(o), 0, 0));
    const T out = T(double(bc) - acc);
    store(y_r, out);
    v[0] += double(bc) * double(bc);
    v[1] += ok ? double(out) * double(out) : 0.0;
    store(ym_r, xc);
    store(sm_r, out);
  }
}

template <class T, int SP, int MODE, int D>
__device__ inline void stl_sweep(const StcView& A, int ic, int jc, int kc, int s, int g, int wv,
                                 int l16, int comp, double w, bool run,
                                 __amdgpu_buffer_rsrc_t xr, __amdgpu_buffer_rsrc_t yr_,
                                 __amdgpu_buffer_rsrc_t br_, __amdgpu_buffer_rsrc_t ymr,
                                 __amdgpu_buffer_rsrc_t smr, typename Vec2<T>::type* pl,
                                 double (&v)[2]) {
  using T2 = typename Vec2<T>::type;
  // 16 lanes per row: 16 systems, or (SPLIT) 8 systems x (re, im), each lane one component
  constexpr bool SPLIT = SP == 8;
  constexpr int NC = kStcCols[D];
  constexpr int NCR = (kStcCoef + 15) / 16;
  constexpr int LO = D == 2 ? 0 : -1, HI = 1;
  constexpr int PS = kStlMaxVal * SP;              // elements of one plane buffer
  const GridDims G{A.nx, A.ny, A.nz};
  const int nxd = G.nxd(D), nyd = G.nyd(D), nzd = G.nzd(D);
  const int i0 = ic * kStlTI, j0 = jc * kStlNW;
  const int k0 = kc * A.kc, k1 = min(k0 + A.kc, nzd);
  if (j0 >= nyd || k0 >= nzd || i0 >= nxd) return;  // the whole tile is empty (uniform)
  const int i = i0 + g, j = j0 + wv;
  const bool row_ok = i < nxd && j < nyd;
  const int ir = min(i, nxd - 1), jr = min(j, nyd - 1);
  // this thread's share of a plane, per column direction dt (so that dt is a compile-time
  // constant of each load): element q = t + kBlock*m of dt's rectangle is system q % SP of its
  // value q / SP; the last load of each dt partly empty (offset kOob: it returns 0 and its
  // value is not stored)
  const int t = int(threadIdx.x);
  constexpr int NL0 = (stl_wid(D, 0) * stl_hgt(D, 0) * SP + kBlock - 1) / kBlock;
  constexpr int NL1 = (stl_wid(D, 1) * stl_hgt(D, 1) * SP + kBlock - 1) / kBlock;
  constexpr int NL2 = (stl_wid(D, 2) * stl_hgt(D, 2) * SP + kBlock - 1) / kBlock;
  constexpr int NL = NL0 + NL1 + NL2;
  constexpr int kFirst[3] = {0, NL0, NL0 + NL1};
  unsigned lo[NL];
  static_for<0, NL>([&](auto Mc) {
    constexpr int m = decltype(Mc)::value;
    constexpr int dt = m < NL0 ? 0 : m < NL0 + NL1 ? 1 : 2;
    constexpr int n = stl_wid(D, dt) * stl_hgt(D, dt) * SP;
    const int q = t + kBlock * (m - kFirst[dt]);
    const int val = min(q, n - 1) / SP, sys = q % SP;
    const int gi = min(max(i0 + stl_rng(D, dt, 0) + val % stl_wid(D, dt), 0), G.nxd(dt) - 1);
    const int gj = min(max(j0 + stl_rng(D, dt, 2) + val / stl_wid(D, dt), 0), G.nyd(dt) - 1);
    const long long row = G.base(dt) + (long long)gj * G.nxd(dt) + gi;
    lo[m] = q < n ? unsigned(row * SP + sys) * unsigned(sizeof(T2)) : kOob;
  });
  auto fetch = [&](T2 (&dst)[NL], int L) {
    // level offsets of the three boxes (uniform)
    const unsigned lv[3] = {
        unsigned(min(max(L, 0), G.nzd(0) - 1)) *
            unsigned((long long)G.nyd(0) * G.nxd(0) * SP * sizeof(T2)),
        unsigned(min(max(L, 0), G.nzd(1) - 1)) *
            unsigned((long long)G.nyd(1) * G.nxd(1) * SP * sizeof(T2)),
        unsigned(min(max(L, 0), G.nzd(2) - 1)) *
            unsigned((long long)G.nyd(2) * G.nxd(2) * SP * sizeof(T2))};
    static_for<0, NL>([&](auto Mc) {
      constexpr int m = decltype(Mc)::value;
      constexpr int dt = m < NL0 ? 0 : m < NL0 + NL1 ? 1 : 2;
      dst[m] = buffer_gather_so<T2>(xr, lo[m], lv[dt]);
    });
  };
  auto put = [&](const T2 (&src)[NL], T2* buf) {
    static_for<0, NL>([&](auto Mc) {
      constexpr int m = decltype(Mc)::value;
      constexpr int dt = m < NL0 ? 0 : m < NL0 + NL1 ? 1 : 2;
      constexpr int n = stl_wid(D, dt) * stl_hgt(D, dt) * SP;
      const int q = t + kBlock * (m - kFirst[dt]);
      if (m - kFirst[dt] < (dt == 0 ? NL0 : dt == 1 ? NL1 : NL2) - 1 || q < n)
        buf[stl_base(D, dt) * SP + q] = src[m];
    });
  };
  // LDS element of column C for this lane: value base(dt) + (wv + dj - jlo) * wid + (g + di -
  // ilo), system s; with lane = g * SP + s that is lane + wv * wid * SP + a constant
  const int lane = g * SP + s;
  const int pw[3] = {lane + wv * stl_wid(D, 0) * SP, lane + wv * stl_wid(D, 1) * SP,
                     lane + wv * stl_wid(D, 2) * SP};
  auto lds_at = [&](auto Cc, const T2* buf) {
    constexpr int C = decltype(Cc)::value;
    constexpr StencilCol c = kStc[D][C];
    constexpr int k = (stl_base(D, c.dt) + (c.dj - stl_rng(D, c.dt, 2)) * stl_wid(D, c.dt) +
                       (c.di - stl_rng(D, c.dt, 0))) * SP;
    return buf[pw[c.dt] + k];
  };
  // the coefficients of direction D's box through a resource of their own, addressed from
  // the box's first row: 32-bit offsets then reach tables beyond 4 GiB (each box under it)
  const __amdgpu_buffer_rsrc_t cb = buf_rsrc(A.coef + G.base(D) * kStcCoef,
                                             G.count(D) * kStcCoef * 8);
  auto load_coef = [&](double (&dst)[NCR], int k) {
    const int kk = min(max(k, 0), nzd - 1);
    const long long r = ((long long)kk * nyd + jr) * nxd + ir;  // row within the box
    const unsigned rb = unsigned(r) * unsigned(kStcCoef * 8);
    static_for<0, NCR>([&](auto Q) {
      const int slot = Q * 16 + l16;
      dst[Q] = buf_f64(cb, slot < kStcCoef ? rb + unsigned(slot) * 8u : kOob);
    });
  };
  constexpr int NBC = 4;
  constexpr int NB = (NC + NBC - 1) / NBC;
  double cf[4][NCR];
  double ar[4], ai[4];
  T2 xs[2];  // the own column's value at the last two levels (the stored row's x)
  xs[0].x = xs[0].y = xs[1].x = xs[1].y = T(0);
  const int L0 = k0 + LO, L1 = k1 - 1 + HI;
  {
    T2 first[NL];
    fetch(first, L0);
    put(first, pl);  // plane L0 -> buffer 0; the first level's barrier publishes it
  }
  load_coef(cf[3], L0 - 1);
  load_coef(cf[0], L0);
  load_coef(cf[1], L0 + 1);
  ar[3] = ai[3] = ar[0] = ai[0] = 0.0;
  // level L = L0 + n, P = n mod 4; plane L in buffer n mod 2
  auto level = [&](auto Pc, int L) {
    constexpr int P = decltype(Pc)::value;
    const T2* cur = pl + (P % 2) * PS;
    T2* nxt_buf = pl + ((P + 1) % 2) * PS;
    T2 nxt[NL];
    const bool more = L + 1 <= L1;
    if (more) fetch(nxt, L + 1);
    load_coef(cf[(P + 2) % 4], L + 2);
    ar[(P + 1) % 4] = ai[(P + 1) % 4] = 0.0;
    __syncthreads();  // plane L is in `cur`; every wave is done with level L-1's buffer
    T2 xb[2][NBC];
    auto read_batch = [&](T2 (&dst)[NBC], auto Bc) {
      constexpr int B = decltype(Bc)::value;
      static_for<0, NBC>([&](auto Jc) {
        constexpr int C = B * NBC + decltype(Jc)::value;
        if constexpr (C < NC) dst[decltype(Jc)::value] = lds_at(std::integral_constant<int, C>{}, cur);
      });
    };
    read_batch(xb[0], std::integral_constant<int, 0>{});
    static_for<0, NB>([&](auto Bc) {
      constexpr int B = decltype(Bc)::value;
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (B + 1 < NB) read_batch(xb[(B + 1) % 2], std::integral_constant<int, B + 1>{});
      __builtin_amdgcn_sched_barrier(0);
      static_for<0, NBC>([&](auto Jc) {
        constexpr int J = decltype(Jc)::value;
        constexpr int C = B * NBC + J;
        if constexpr (C < NC) {
          constexpr StencilCol c = kStc[D][C];
          const double xr_ = double(xb[B % 2][J].x), xi_ = double(xb[B % 2][J].y);
          const double nwxi = c.dt == D ? -w * xi_ : 0.0, wxr = c.dt == D ? w * xr_ : 0.0;
          // SPLIT: this lane's component of x and of the i w b term
          const double u = comp ? xi_ : xr_, vv = comp ? wxr : nwxi;
          static_for<c.lo, c.hi + 1>([&](auto DK) {
            constexpr int dk = decltype(DK)::value;
            constexpr int R = ((P - dk) % 4 + 4) % 4;
            constexpr int o = stc_slot(D, C, dk);
            if constexpr (SPLIT) {
              if constexpr (c.dt == D)
                dfma2x_bc<o % 16, (o + 1) % 16>(ar[R], cf[R][o / 16], cf[R][(o + 1) / 16], u,
                                                vv);
              else
                dfma1_bc<o % 16>(ar[R], cf[R][o / 16], u);
            } else {
              if constexpr (c.dt == D)
                dfma4_bc<o % 16, (o + 1) % 16>(ar[R], ai[R], cf[R][o / 16], cf[R][(o + 1) / 16],
                                               xr_, xi_, nwxi, wxr);
              else
                dfma2_bc<o % 16>(ar[R], ai[R], cf[R][o / 16], xr_, xi_);
            }
          });
        }
      });
    });
    // row k = L - 1 (HI = 1 for every direction) is complete; its own x is the own column's
    // value at level L - 1, kept from that level
    static_assert(HI == 1, "rows complete one level after their own");
    xs[P % 2] = lds_at(std::integral_constant<int, stl_self(D)>{}, cur);
    const int k = L - HI;
    constexpr int R = ((P - HI) % 4 + 4) % 4;
    const bool ok = k >= k0 && k < k1 && row_ok && (run || MODE == 0);
    const long long tt =
        (G.base(D) + ((long long)min(max(k, 0), nzd - 1) * nyd + jr) * nxd + ir) * SP + s;
    if constexpr (SPLIT)
      stl_store_split<T, MODE>(ok, unsigned(tt) * unsigned(sizeof(T2)), comp, ar[R],
                               xs[(P + 1) % 2], yr_, br_, ymr, smr, v);
    else
      stl_store<T, MODE>(ok, unsigned(tt) * unsigned(sizeof(T2)), ar[R], ai[R],
                         xs[(P + 1) % 2], yr_, br_, ymr, smr, v);
    if (more) put(nxt, nxt_buf);
  };
  for (int L = L0; L <= L1; L += 4) {
    level(std::integral_constant<int, 0>{}, L);
    if (L + 1 > L1) break;
    level(std::integral_constant<int, 1>{}, L + 1);
    if (L + 2 > L1) break;
    level(std::integral_constant<int, 2>{}, L + 2);
    if (L + 3 > L1) break;
    level(std::integral_constant<int, 3>{}, L + 3);
  }
  __syncthreads();  // the last plane's readers are done before the next tile writes buffer 0
}

template <class T, int SP, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_stl(StcView A, const double* omega,
                                                     const typename Vec2<T>::type* __restrict__ x,
                                                     typename Vec2<T>::type* __restrict__ y,
                                                     const typename Vec2<T>::type* __restrict__ b,
                                                     typename Vec2<T>::type* __restrict__ ym,
                                                     typename Vec2<T>::type* __restrict__ sm,
                                                     const SysScalars* __restrict__ sc,
                                                     const int* __restrict__ nrun,
                                                     double* partials) {
  // 16-lane rows (the DPP row broadcast): 16 systems, or 8 systems x (re, im)
  static_assert(SP == 16 || SP == 8, "16 or 8 systems");
  using T2 = typename Vec2<T>::type;
  __shared__ T2 pl[2 * kStlMaxVal * SP];
  if (*nrun == 0) return;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int l16 = lane & 15, g = lane >> 4;
  const int s = SP == 16 ? l16 : l16 >> 1, comp = SP == 16 ? 0 : l16 & 1;
  const bool run = (MODE == 2 ? sc[s].outer_status : sc[s].status) == ST_RUNNING;
  const double w = omega[s];
  double v[2] = {0.0, 0.0};
  const long long vbytes = (long long)A.nb * SP * (long long)sizeof(T2);
  const __amdgpu_buffer_rsrc_t xr = buf_rsrc(x, vbytes);
  const __amdgpu_buffer_rsrc_t y_r = buf_rsrc(y, vbytes);
  const __amdgpu_buffer_rsrc_t b_r = buf_rsrc(b, b ? vbytes : 0);
  const __amdgpu_buffer_rsrc_t ym_r = buf_rsrc(ym, ym ? vbytes : 0);
  const __amdgpu_buffer_rsrc_t sm_r = buf_rsrc(sm, sm ? vbytes : 0);
  // tasks (sweep chunk, band of nbw tile columns, tile row jc, tile column within the band,
  // direction), direction fastest; one per workgroup in the XCD-aware walk (every wave of a
  // workgroup takes the same task). The workgroups of one XCD then work on a band nbw tiles
  // wide, so that the tiles that share an x halo run close in time, within one L2
  TileWalk tw(A.ntasks);
  const long long per_kc = (long long)A.njc * A.nic;
  for (long long task = tw.first; task < tw.last; task += tw.step) {
    const int d = int(task % 3);
    const long long q = task / 3;
    const int kc = int(q / per_kc);
    const long long r = q % per_kc;
    const long long full = (long long)A.njc * A.nbw;  // tasks of a full band
    const int band = int(r / full);
    const int wb = min(A.nbw, A.nic - band * A.nbw);  // the last band may be narrower
    const long long rr = r - band * full;
    const int jc = int(rr / wb);
    const int ic = band * A.nbw + int(rr % wb);
    if (d == 0)
      stl_sweep<T, SP, MODE, 0>(A, ic, jc, kc, s, g, wv, l16, comp, w, run, xr, y_r, b_r,
                                ym_r, sm_r, pl, v);
    else if (d == 1)
      stl_sweep<T, SP, MODE, 1>(A, ic, jc, kc, s, g, wv, l16, comp, w, run, xr, y_r, b_r,
                                ym_r, sm_r, pl, v);
    else
      stl_sweep<T, SP, MODE, 2>(A, ic, jc, kc, s, g, wv, l16, comp, w, run, xr, y_r, b_r,
                                ym_r, sm_r, pl, v);
  }
  if constexpr (SP == 8) {
    v[0] += __shfl_xor(v[0], 1, 64);  // the (re, im) lane pair of a system
    v[1] += __shfl_xor(v[1], 1, 64);
    block_partials_split<SP, 2>(v, partials);
  } else {
    block_partials<SP, 2>(v, partials);
  }
}

// The folded preconditioner's arithmetic, shared by the node and edge kernels and the fused
// grid kernel so that every path gives the same bits: G and G^T products accumulate with fused
// multiply-adds in entry order, complex products round each term.
template <class T>
__device__ inline T fmad(T a, T b, T c);
template <>
__device__ inline double fmad<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ inline float fmad<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <class T, class T2>
__device__ inline void pc_acc(T& yr, T& yi, T g, T2 v) {
  yr = fmad<T>(g, v.x, yr);
  yi = fmad<T>(g, v.y, yi);
}
template <class T, class T2>
__device__ inline void pc_cmul(T ar, T ai, T2 d, T& yr, T& yi) {
#pragma clang fp contract(off)
  yr = ar * d.x - ai * d.y;
  yi = ar * d.y + ai * d.x;
}

// help = D_N^-1 (G v), v = z + beta*a (USE_BETA) or v = z  (FoldedPreconditioner.cpp:598-615)
template <class T, int SP, bool USE_BETA>
__global__ __launch_bounds__(kBlock) void k_pc_node(GView<T> G,
                                                    const typename Vec2<T>::type* __restrict__ z,
                                                    const typename Vec2<T>::type* __restrict__ a,
                                                    const typename Vec2<T>::type* __restrict__ dn,
                                                    typename Vec2<T>::type* __restrict__ help,
                                                    const SysScalars* __restrict__ sc,
                                                    const int* __restrict__ nrun) {
  using T2 = typename Vec2<T>::type;
  if (*nrun == 0) return;
  const long long N = (long long)G.m * SP;
  const long long ntiles = (N + kBlock - 1) / kBlock;
  TileWalk tw(ntiles);
  // kBlock is a multiple of SP: a thread serves one system in every tile
  const int s = int(threadIdx.x % SP);
  // lanes of a system not running work on with beta = 0 (their help is dead and stays
  // bounded), so that the wave's memory instructions keep every lane active
  const bool run = sc[s].status == ST_RUNNING;
  const double br = USE_BETA && run ? sc[s].be_re : 0.0, bi = USE_BETA && run ? sc[s].be_im : 0.0;
  for (long long tile = tw.first; tile < tw.last; tile += tw.step) {
    const long long t = tile * kBlock + threadIdx.x;
    if (t >= N) continue;
    const int i = int(t / SP);
    T hr = 0, hi = 0;
    // the first kGU entries (a node of a hexahedral mesh has at most 6 edges) with every load
    // issued before the arithmetic, the rest (other meshes) one by one; the sum keeps the order
    constexpr int kGU = 6;
    int ec[kGU];
    T gv[kGU];
    T2 zv[kGU], av[kGU];
#pragma unroll
    for (int k = 0; k < kGU; ++k) {
      ec[k] = k < G.wg ? G.gcol[(size_t)k * G.m + i] : 0;
      gv[k] = k < G.wg ? G.gval[(size_t)k * G.m + i] : T(0);
    }
#pragma unroll
    for (int k = 0; k < kGU; ++k) {
      if (k < G.wg) {
        zv[k] = z[(size_t)ec[k] * SP + s];
        if (USE_BETA) av[k] = a[(size_t)ec[k] * SP + s];
      }
    }
#pragma unroll
    for (int k = 0; k < kGU; ++k) {
      if (k < G.wg) {
        T2 vv = zv[k];
        if (USE_BETA) vv = caxpy(vv, av[k], br, bi);
        pc_acc(hr, hi, gv[k], vv);
      }
    }
    for (int k = kGU; k < G.wg; ++k) {
      const int e = G.gcol[(size_t)k * G.m + i];
      const T g = G.gval[(size_t)k * G.m + i];
      T2 vv = z[(size_t)e * SP + s];
      if (USE_BETA) vv = caxpy(vv, a[(size_t)e * SP + s], br, bi);
      pc_acc(hr, hi, g, vv);
    }
    T2 o;
    pc_cmul(hr, hi, dn[t], o.x, o.y);
    help[t] = o;
  }
}

// y += (G^T help)_e for one system: the first two entries (an edge has two end nodes) with
// their loads issued together, any further ones (other meshes) one by one, in entry order
template <class T, int SP>
__device__ inline void gt_apply(const GView<T>& G, const typename Vec2<T>::type* __restrict__ help,
                                int e, int s, T& yr, T& yi) {
  using T2 = typename Vec2<T>::type;
  constexpr int kGU = 2;
  int nc[kGU];
  T gv[kGU];
  T2 hv[kGU];
#pragma unroll
  for (int k = 0; k < kGU; ++k) {
    nc[k] = k < G.wgt ? G.gtcol[(size_t)k * G.n + e] : 0;
    gv[k] = k < G.wgt ? G.gtval[(size_t)k * G.n + e] : T(0);
  }
#pragma unroll
  for (int k = 0; k < kGU; ++k)
    if (k < G.wgt) hv[k] = help[(size_t)nc[k] * SP + s];
#pragma unroll
  for (int k = 0; k < kGU; ++k)
    if (k < G.wgt) pc_acc(yr, yi, gv[k], hv[k]);
  for (int k = kGU; k < G.wgt; ++k) {
    const int node = G.gtcol[(size_t)k * G.n + e];
    const T g = G.gtval[(size_t)k * G.n + e];
    pc_acc(yr, yi, g, help[(size_t)node * SP + s]);
  }
}

// Prologue form: out = D_A^-1 v + G^T help (FoldedPreconditioner.cpp:585-589,618-627)
template <class T, int SP>
__global__ __launch_bounds__(kBlock) void k_pc_edge_plain(
    GView<T> G, const typename Vec2<T>::type* __restrict__ v, DiagA da,
    const typename Vec2<T>::type* __restrict__ help, typename Vec2<T>::type* __restrict__ out,
    const SysScalars* __restrict__ sc, const int* __restrict__ nrun) {
  using T2 = typename Vec2<T>::type;
  if (*nrun == 0) return;
  const long long N = (long long)G.n * SP;
  const long long ntiles = (N + kBlock - 1) / kBlock;
  TileWalk tw(ntiles);
  const int s = int(threadIdx.x % SP);  // kBlock is a multiple of SP
  if (sc[s].status != ST_RUNNING) return;
  for (long long tile = tw.first; tile < tw.last; tile += tw.step) {
    const long long t = tile * kBlock + threadIdx.x;
    if (t >= N) continue;
    const int e = int(t / SP);
    const T2 xv = v[t];
    T yr, yi;
    pc_cmul(xv.x, xv.y, diag_inv<T>(da, e, s), yr, yi);
    gt_apply<T, SP>(G, help, e, s, yr, yi);
    T2 o;
    o.x = yr;
    o.y = yi;
    out[t] = o;
  }
}

// Iteration form (stages 5-6): a = z + beta*a, w = M^-1 a, partial (a, w). (p = s + beta*p of
// the same stage is formed by k_xr_fused, which reads s and p anyway.)
template <class T, int SP>
__global__ __launch_bounds__(kBlock) void k_pc_edge_iter(
    GView<T> G, const typename Vec2<T>::type* __restrict__ z, typename Vec2<T>::type* __restrict__ a,
    typename Vec2<T>::type* __restrict__ w, DiagA da,
    const typename Vec2<T>::type* __restrict__ help, const SysScalars* __restrict__ sc,
    const int* __restrict__ nrun, double* partials) {
  using T2 = typename Vec2<T>::type;
  if (*nrun == 0) return;
  const long long N = (long long)G.n * SP;
  const long long ntiles = (N + kBlock - 1) / kBlock;
  double v[2] = {0.0, 0.0};
  TileWalk tw(ntiles);
  const int s = int(threadIdx.x % SP);  // kBlock is a multiple of SP
  // lanes of a system not running work on with beta = 0 (their a, w are dead and stay
  // bounded), so that the wave's memory instructions keep every lane active
  const bool run = sc[s].status == ST_RUNNING;
  const double br = run ? sc[s].be_re : 0.0, bi = run ? sc[s].be_im : 0.0;
  // two elements per thread and step, streams nontemporal (help is gathered and stays cached)
  auto elem = [&](long long t, T2 zv, T2 av) {
    const int e = int(t / SP);
    const T2 an = caxpy(zv, av, br, bi);
    T yr, yi;
    pc_cmul(an.x, an.y, diag_inv<T>(da, e, s), yr, yi);
    gt_apply<T, SP>(G, help, e, s, yr, yi);
    T2 o;
    o.x = yr;
    o.y = yi;
    st_nt(a + t, an);
    st_nt(w + t, o);
    v[0] += double(an.x) * double(yr) - double(an.y) * double(yi);
    v[1] += double(an.x) * double(yi) + double(an.y) * double(yr);
  };
  for (long long tile = tw.first; tile < tw.last; tile += 2 * tw.step) {
    const long long t0 = tile * kBlock + threadIdx.x;
    const long long t1 = (tile + tw.step) * kBlock + threadIdx.x;
    const bool ok0 = t0 < N;
    const bool ok1 = tile + tw.step < tw.last && t1 < N;
    T2 z0, a0, z1, a1;
    if (ok0) {
      z0 = ld_nt(z + t0);
      a0 = ld_nt(a + t0);
    }
    if (ok1) {
      z1 = ld_nt(z + t1);
      a1 = ld_nt(a + t1);
    }
    if (ok0) elem(t0, z0, a0);
    if (ok1) elem(t1, z1, a1);
  }
  block_partials<SP, 2>(v, partials);
}

// Fused preconditioner pass on a box mesh held in the grid order (8 or 16 systems): the node pass
// (help = D_N^-1 G a) and the edge pass (a = z + beta a, w = D_A^-1 a + G^T help, partial (a, w))
// of one iteration in one kernel, with help kept in LDS instead of written to and re-read from
// HBM, and each edge's z and a read once instead of three times. A tile reads the a of edges
// that neighbouring tiles own, so the new a goes to the other of two buffers (`ain` holds the
// previous iteration's, `aout` receives this one's; they alternate with the iterations).
//
// Every node (i, j, k) owns its forward edges: the x-edge to (i+1, j, k), the y-edge to
// (i, j+1, k) and the z-edge to (i, j, k+1), where they exist. A workgroup takes a tile of
// kPgTI x kPgTJ nodes in (i, j) and sweeps a chunk of node levels in k. At level L it forms a
// for the edges it needs, help(L) of the tile's nodes and of the halo nodes one step beyond in
// +i and +j (computed again by the neighbouring tile) into an LDS plane, then w of the tile's
// x- and y-edges at level L and of its z-edges at level L - 1 (whose upper node's help is
// help(L)). Three planes rotate, so one barrier per level suffices: the plane a level writes
// was last read two levels before, behind the previous level's barrier. An 8 x 4 tile (two
// owned nodes per slot, 164 VGPRs, three workgroups per CU) ran 2% faster per solve than 8 x 8
// (four nodes per slot, two workgroups per CU), and chunks of 8 levels faster than 4, 6, 12
// or 16. The z-edge below a node is the z-edge above it one level down: its a is kept in a
// register from the previous level. G's rows of a box mesh are the six edges of an interior node
// in ascending reference edge order (z below, y below, x below, x above, y above, z above, the
// first three +1/2, the others -1/2), G^T's rows the edge's lower then upper node (-1/2, +1/2);
// the host checks the uploaded G against this before the kernel is used, so the sums run in
// the node and edge kernels' order and give their bits.
constexpr int kPgTI = 8, kPgTJ = 4;            // tile sides in nodes (i, j)
constexpr int kPgE = kPgTI + 1;                // row length of the tile with its halo
constexpr int kPgEJ = kPgTJ + 1;               // rows of the tile with its halo
constexpr int kPgNH = kPgTJ + 1 + kPgTI;          // halo nodes: the column i0 + TI, the row j0 + TJ
// per batch width SP: node slots of a workgroup (SP lanes each), owned and halo nodes per slot
template <int SP>
constexpr int pg_slots() { return kBlock / SP; }
template <int SP>
constexpr int pg_own() { return kPgTI * kPgTJ / pg_slots<SP>(); }
template <int SP>
constexpr int pg_halo() { return (kPgNH + pg_slots<SP>() - 1) / pg_slots<SP>(); }
// the batch widths the fused kernel serves (a whole number of owned nodes per slot)
constexpr bool pg_width(int sp) { return sp == 8 || sp == 16; }

struct PgView {
  int nx, ny, nz;      // nodes per axis of the box
  int nti, ntj;        // tiles in i and j
  int kc, nkc;         // node levels per chunk, chunks
  long long ntasks;
  int nb;
};

//
// HELP_OUT (the FP64 solver): instead of w, the kernel writes help of the nodes it owns (interior
// nodes of its tile at its owned levels) to `w`, as a node vector [interior node][system]; the
// update pass forms w = D_A^-1 a + G^T help of each edge from it (k_xr_fused<.., .., true,
// true>), with the same arithmetic, so the 16-byte w per edge and system is neither written here
// nor read there (help is 16 bytes per node and system, about a third of an edge's). The partial
// (a, w) is formed here as before.
template <class T, int SP, bool HELP_OUT = false>
__global__ __launch_bounds__(kBlock) void k_pc_grid(
    PgView P, const typename Vec2<T>::type* __restrict__ z,
    const typename Vec2<T>::type* __restrict__ ain, typename Vec2<T>::type* __restrict__ aout,
    typename Vec2<T>::type* __restrict__ w, const typename Vec2<T>::type* __restrict__ dn,
    DiagA da, const SysScalars* __restrict__ sc, const int* __restrict__ nrun,
    double* partials) {
  static_assert(pg_width(SP) && kPgTI * kPgTJ % pg_slots<SP>() == 0,
                "a whole number of owned nodes per slot");
  constexpr int kPgSlots = pg_slots<SP>(), kPgOwn = pg_own<SP>(), kPgHalo = pg_halo<SP>();
  using T2 = typename Vec2<T>::type;
  __shared__ T2 hp[3][kPgE * kPgEJ * SP];
  if (*nrun == 0) return;
  const int s = int(threadIdx.x) % SP, q = int(threadIdx.x) / SP;
  const bool run = sc[s].status == ST_RUNNING;
  const double br = run ? sc[s].be_re : 0.0, bi = run ? sc[s].be_im : 0.0;
  const long long vb = (long long)P.nb * SP * (long long)sizeof(T2);
  const __amdgpu_buffer_rsrc_t zr = buf_rsrc(z, vb), ar = buf_rsrc(ain, vb);
  const long long m = (long long)(P.nx - 2) * (P.ny - 2) * (P.nz - 2);
  const __amdgpu_buffer_rsrc_t aw = buf_rsrc(aout, vb);
  const __amdgpu_buffer_rsrc_t wr = buf_rsrc(w, HELP_OUT ? m * SP * (long long)sizeof(T2) : vb);
  const __amdgpu_buffer_rsrc_t dr = buf_rsrc(dn, m * SP * (long long)sizeof(T2));
  const GridDims G{P.nx, P.ny, P.nz};
  auto store = [&](__amdgpu_buffer_rsrc_t r, unsigned off, T2 val) {
    if constexpr (sizeof(T2) == 16)
      __builtin_amdgcn_raw_buffer_store_b128(
          __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, val), r, int(off), 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b64(
          __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, val), r, int(off), 0, 0);
  };
  // byte offset of system s of edge (d, i, j, k), or kOob where the edge does not exist
  auto eoff = [&](int d, int i, int j, int k) -> unsigned {
    const bool ok = i >= 0 && j >= 0 && k >= 0 && i < G.nxd(d) && j < G.nyd(d) && k < G.nzd(d);
    const long long row = G.base(d) + ((long long)k * G.nyd(d) + j) * G.nxd(d) + i;
    return ok ? unsigned((row * SP + s) * (long long)sizeof(T2)) : kOob;
  };
  auto interior = [&](int i, int j, int k) {
    return i >= 1 && j >= 1 && k >= 1 && i <= P.nx - 2 && j <= P.ny - 2 && k <= P.nz - 2;
  };
  // a = z + beta a of an edge (0 where the edge does not exist: both loads return 0)
  auto anew = [&](unsigned off) {
    return caxpy(buffer_gather<T2>(zr, off), buffer_gather<T2>(ar, off), br, bi);
  };
  // D_A^-1 of an edge (0 where it does not exist). Evaluated for row 0 there and selected away,
  // so that no load sits behind a branch (see help_of).
  auto dinv = [&](int d, int i, int j, int k) {
    const bool ok = i < G.nxd(d) && j < G.nyd(d) && k < G.nzd(d);
    const long long row = ok ? G.base(d) + ((long long)k * G.nyd(d) + j) * G.nxd(d) + i : 0;
    const T2 q = diag_inv<T>(da, int(row), s);
    T2 o;
    o.x = ok ? q.x : T(0);
    o.y = ok ? q.y : T(0);
    return o;
  };
  const T half = T(0.5);
  double v[2] = {0.0, 0.0};
  TileWalk tw(P.ntasks);
  for (long long task = tw.first; task < tw.last; task += tw.step) {
    const int ti = int(task % P.nti);
    const long long r1 = task / P.nti;
    const int tj = int(r1 % P.ntj);
    const int kcI = int(r1 / P.ntj);
    const int i0 = ti * kPgTI, j0 = tj * kPgTJ;
    const int k0 = kcI * P.kc, k1 = min(k0 + P.kc, P.nz);  // owned node levels [k0, k1)
    const int Lend = min(k1, P.nz - 1);                    // help levels k0 .. Lend
    // the slot's nodes: owned o = q + kPgSlots*r of the tile, then halo nodes (the column
    // i = i0 + kPgTI, then the row j = j0 + kPgTJ)
    int oi[kPgOwn], oj[kPgOwn], hi_[kPgHalo], hj_[kPgHalo];
    bool hv[kPgHalo];
#pragma unroll
    for (int r = 0; r < kPgOwn; ++r) {
      const int o = q + kPgSlots * r;
      oi[r] = i0 + o % kPgTI;
      oj[r] = j0 + o / kPgTI;
    }
#pragma unroll
    for (int r = 0; r < kPgHalo; ++r) {
      const int hh = q + kPgSlots * r;
      hv[r] = hh < kPgNH;
      hi_[r] = hh <= kPgTJ ? i0 + kPgTI : i0 + (hh - kPgTJ - 1);
      hj_[r] = hh <= kPgTJ ? j0 + hh : j0 + kPgTJ;
    }
    // a of the z-edge below each node at the first level (kept from level to level after)
    T2 zo[kPgOwn], zh[kPgHalo];
#pragma unroll
    for (int r = 0; r < kPgOwn; ++r) zo[r] = anew(eoff(2, oi[r], oj[r], k0 - 1));
#pragma unroll
    for (int r = 0; r < kPgHalo; ++r) zh[r] = anew(hv[r] ? eoff(2, hi_[r], hj_[r], k0 - 1) : kOob);
    for (int L = k0; L <= Lend; ++L) {
      T2* cur = hp[L % 3];
      const T2* prv = hp[(L + 2) % 3];
      // help of a node at level L from the a of its six edges (zb: the z-edge below), 0 for a
      // node on the boundary or a slot without a halo node (valid false). Its loads run
      // whatever the node is (with offset kOob where they must not touch memory: they then
      // return 0) and the result is selected, so that no load sits behind a branch, which
      // would make it wait for the loads before it (on config 2: 1.2% off a batched step).
      auto help_of = [&](bool valid, int i, int j, T2 zb, T2 xp, T2 yp, T2 zp) {
        const bool in = valid && interior(i, j, L);
        const T2 xm = anew(in ? eoff(0, i - 1, j, L) : kOob);
        const T2 ym = anew(in ? eoff(1, i, j - 1, L) : kOob);
        const long long mi =
            (i - 1) + (long long)(P.nx - 2) * ((j - 1) + (long long)(P.ny - 2) * (L - 1));
        const T2 dv =
            buffer_gather<T2>(dr, in ? unsigned((mi * SP + s) * (long long)sizeof(T2)) : kOob);
        T hr = T(0), hi = T(0);
        pc_acc(hr, hi, half, zb);
        pc_acc(hr, hi, half, ym);
        pc_acc(hr, hi, half, xm);
        pc_acc(hr, hi, -half, xp);
        pc_acc(hr, hi, -half, yp);
        pc_acc(hr, hi, -half, zp);
        T2 c;
        pc_cmul(hr, hi, dv, c.x, c.y);
        T2 o;
        o.x = in ? c.x : T(0);
        o.y = in ? c.y : T(0);
        return o;
      };
      T2 xo[kPgOwn], yo[kPgOwn], zn[kPgOwn];
#pragma unroll
      for (int r = 0; r < kPgOwn; ++r) {
        xo[r] = anew(eoff(0, oi[r], oj[r], L));
        yo[r] = anew(eoff(1, oi[r], oj[r], L));
        zn[r] = anew(eoff(2, oi[r], oj[r], L));
        cur[((oj[r] - j0) * kPgE + (oi[r] - i0)) * SP + s] =
            help_of(true, oi[r], oj[r], zo[r], xo[r], yo[r], zn[r]);
      }
#pragma unroll
      for (int r = 0; r < kPgHalo; ++r) {
        // a slot without a halo node (hv false) loads nothing (kOob) and stores nothing
        const T2 xh = anew(hv[r] ? eoff(0, hi_[r], hj_[r], L) : kOob);
        const T2 yh = anew(hv[r] ? eoff(1, hi_[r], hj_[r], L) : kOob);
        const T2 zhn = anew(hv[r] ? eoff(2, hi_[r], hj_[r], L) : kOob);
        const T2 hh = help_of(hv[r], hi_[r], hj_[r], zh[r], xh, yh, zhn);
        if (hv[r]) cur[((hj_[r] - j0) * kPgE + (hi_[r] - i0)) * SP + s] = hh;
        zh[r] = zhn;
      }
      __syncthreads();  // help(L) of the tile and its halo is in `cur`
      // w = D_A^-1 a + G^T help of the owned edges: x and y at level L (if L is owned), z at
      // level L - 1 (its upper node is at L)
#pragma unroll
      for (int r = 0; r < kPgOwn; ++r) {
        const int i = oi[r], j = oj[r];
        const int c = ((j - j0) * kPgE + (i - i0)) * SP + s;
        const bool own_n = interior(i, j, L);
        const T2 hn = cur[c];
        auto finish = [&](unsigned off, T2 an, T2 d, bool lo_in, T2 hlo, bool up_in, T2 hup) {
          T yr, yi;
          pc_cmul(an.x, an.y, d, yr, yi);
          if (lo_in) pc_acc(yr, yi, -half, hlo);
          if (up_in) pc_acc(yr, yi, half, hup);
          T2 o;
          o.x = yr;
          o.y = yi;
          store(aw, off, an);
          if (!HELP_OUT) store(wr, off, o);
          if (off != kOob) {
            v[0] += double(an.x) * double(yr) - double(an.y) * double(yi);
            v[1] += double(an.x) * double(yi) + double(an.y) * double(yr);
          }
        };
        if (L < k1) {
          finish(eoff(0, i, j, L), xo[r], dinv(0, i, j, L), own_n, hn, interior(i + 1, j, L),
                 cur[c + SP]);
          finish(eoff(1, i, j, L), yo[r], dinv(1, i, j, L), own_n, hn, interior(i, j + 1, L),
                 cur[c + kPgE * SP]);
        }
        if (L > k0)
          finish(eoff(2, i, j, L - 1), zo[r], dinv(2, i, j, L - 1), interior(i, j, L - 1), prv[c],
                 own_n, hn);
        if constexpr (HELP_OUT) {
          // help of the node itself at an owned level (the level above the chunk is the next
          // chunk's)
          const long long mi =
              (i - 1) + (long long)(P.nx - 2) * ((j - 1) + (long long)(P.ny - 2) * (L - 1));
          store(wr, own_n && L < k1 ? unsigned((mi * SP + s) * (long long)sizeof(T2)) : kOob, hn);
        }
        zo[r] = zn[r];
      }
    }
    __syncthreads();  // the task's last planes are read before the next task writes its first
  }
  block_partials<SP, 2>(v, partials);
}

// One pass per iteration k over the edge vectors (pcocr_rci.cpp:139-154, MRS.cpp:85-153):
//   1. the MRS step of iteration k - 1, whose weight s_mrs_stop computed from the previous
//      pass: ym += w (x - ym), sm += w (r - sm) (or a copy, or nothing), partial ||sm||^2 for
//      the deferred stop test of k - 1;
//   2. iteration k: p = s + beta p (stage 5), x += alpha p (into the other x buffer, so that x of
//      k - 1 survives if k - 1 turns out to be the last iteration), r -= alpha a, s -= alpha w,
//      and the partials (sm, u), (u, u) of u = r - sm for the MRS weight of k.
// Step 2 is skipped by a system that halts in iteration k. Deferring the MRS step saves
// re-reading x and r, and forming p here saves the preconditioner pass reading s and p: 8 vector
// reads and 6 writes instead of 13 and 6. With ym null (the FP64 solver, which returns x
// unsmoothed) the ym blend is skipped: 7 reads and 5 writes.
//
// LAZY (the FP64 solver): x and p are left out (k_flush_xp applies their recurrence four
// iterations at a time), and s_k = s_{k-1} - alpha w goes to the next buffer of a ring
// (s_src -> s_vec) instead of overwriting s_{k-1}, which the x/p recurrence still needs. The
// pass then moves 8 vectors instead of 12. Eager mode passes s_src == s_vec (each element is
// read before it is written, by the same thread).
//
// RECW (with LAZY, after k_pc_grid<.., .., true>): w is not read but formed here, as
// w = D_A^-1 a + G^T help (FoldedPreconditioner.cpp:616-629) from this edge's a, its diagonal
// and the help of its two end nodes, which `w` then points to (a node vector
// [interior node][system], `wrec` its box): the arithmetic of the fused preconditioner's edge
// step in the same order, so the same bits.
struct WRec {
  int nx = 0, ny = 0, nz = 0;  // nodes per axis of the box (the matrix in the grid order)
  DiagA da{};
};

template <class T, int SP, bool LAZY = false, bool RECW = false>
__global__ __launch_bounds__(kBlock) void k_xr_fused(
    const typename Vec2<T>::type* __restrict__ xin, typename Vec2<T>::type* __restrict__ xout,
    typename Vec2<T>::type* __restrict__ p, typename Vec2<T>::type* __restrict__ r,
    const typename Vec2<T>::type* __restrict__ a, const typename Vec2<T>::type* s_src,
    typename Vec2<T>::type* s_vec, const typename Vec2<T>::type* __restrict__ w,
    typename Vec2<T>::type* __restrict__ ym, typename Vec2<T>::type* __restrict__ sm,
    const SysScalars* __restrict__ sc, const int* __restrict__ nrun, double* partials, int nb,
    WRec wrec = WRec{}) {
  using T2 = typename Vec2<T>::type;
  static_assert(!RECW || LAZY, "w is formed in the pass of the FP64 solver only");
  if (*nrun == 0) return;
  const long long N = (long long)nb * SP;
  const long long ntiles = (N + kBlock - 1) / kBlock;
  double v[3] = {0.0, 0.0, 0.0};
  TileWalk tw(ntiles);
  // kBlock is a multiple of SP: a thread serves one system in every tile, so the system's
  // scalars are read once and every branch below is the same for all of the thread's tiles
  const int s = int(threadIdx.x % SP);
  // A lane whose system is not running (retired, or padding of the batch) takes the running
  // lanes' path with zero weight and zero alpha and beta, and reads x from x_out and stores it
  // back unchanged (x_out may hold the answer). The memory instructions of a wave then run
  // with every lane active: with idle lanes masked off, this pass ran 50% slower (0.81 vs
  // 0.53 ms at 8 systems on a 1.84 M-edge mesh).
  // With every slot of the batch running (the common case) the pass is compiled without the
  // idle-lane selects.
  auto pass = [&](auto all_running) {
    const bool run = decltype(all_running)::value || sc[s].status == ST_RUNNING;
    const int mode = run ? sc[s].mode : MRS_BLEND;
    const T wm = run ? T(sc[s].w) : T(0);
    const bool step2 = !run || !sc[s].halt;
    const double ar = run ? sc[s].al_re : 0.0, ai = run ? sc[s].al_im : 0.0;
    const double br = run ? sc[s].be_re : 0.0, bi = run ? sc[s].be_im : 0.0;
    const typename Vec2<T>::type* const xsrc = run ? xin : xout;
    // Two elements per thread and step (tiles tile and tile + step), every load of both issued
    // before any use, nontemporal loads and stores: the MRS pass runs 12 streams at once, and
    // this form keeps more bytes in flight per wave.
    struct Elem {
      T2 x, r, sm, s, p, a, w;
      T2 hlo, hup, dinv;  // RECW: help of the edge's end nodes, D_A^-1
      bool lo_in, up_in;
    };
    // RECW: the edge's box (direction d) and position, its end nodes and their help
    const GridDims G{wrec.nx, wrec.ny, wrec.nz};
    const long long mnodes = (long long)(wrec.nx - 2) * (wrec.ny - 2) * (wrec.nz - 2);
    const __amdgpu_buffer_rsrc_t hr =
        buf_rsrc(RECW ? w : nullptr, RECW ? mnodes * SP * (long long)sizeof(T2) : 0);
    auto interior = [&](int i, int j, int k) {
      return i >= 1 && j >= 1 && k >= 1 && i <= G.nx - 2 && j <= G.ny - 2 && k <= G.nz - 2;
    };
    auto hoff = [&](bool in, int i, int j, int k) -> unsigned {
      const long long mi = (i - 1) + (long long)(G.nx - 2) * ((j - 1) + (long long)(G.ny - 2) * (k - 1));
      return in ? unsigned((mi * SP + s) * (long long)sizeof(T2)) : kOob;
    };
    auto load = [&](long long t, Elem& E) {
      if (!LAZY) E.x = ld_nt(xsrc + t);
      E.r = ld_nt(r + t);
      E.sm = ld_nt(sm + t);
      if (step2) {
        E.s = ld_nt(s_src + t);
        if (!LAZY) E.p = ld_nt(p + t);
        E.a = ld_nt(a + t);
        if constexpr (RECW) {
          const long long row = t / SP;
          const long long c0 = G.count(0), c1 = G.count(1);
          const int d = row < c0 ? 0 : row < c0 + c1 ? 1 : 2;
          const long long rr = row - G.base(d);
          const int i = int(rr % G.nxd(d));
          const long long q = rr / G.nxd(d);
          const int j = int(q % G.nyd(d)), k = int(q / G.nyd(d));
          const int iu = i + (d == 0), ju = j + (d == 1), ku = k + (d == 2);
          E.lo_in = interior(i, j, k);
          E.up_in = interior(iu, ju, ku);
          E.hlo = buffer_gather<T2>(hr, hoff(E.lo_in, i, j, k));
          E.hup = buffer_gather<T2>(hr, hoff(E.up_in, iu, ju, ku));
          E.dinv = diag_inv<T>(wrec.da, int(row), s);
        } else {
          E.w = ld_nt(w + t);
        }
      }
    };
    auto update = [&](long long t, Elem& E) {
      if constexpr (RECW) {
        if (step2) {
          // as k_pc_grid's edge step: w = a D_A^-1, - help(lower)/2, + help(upper)/2
          const T half = T(0.5);
          T yr, yi;
          pc_cmul(E.a.x, E.a.y, E.dinv, yr, yi);
          if (E.lo_in) pc_acc(yr, yi, -half, E.hlo);
          if (E.up_in) pc_acc(yr, yi, half, E.hup);
          E.w.x = yr;
          E.w.y = yi;
        }
      }
      T2 smv = E.sm;
      if (mode == MRS_BLEND) {
        if (!LAZY && ym) {
          const T2 y0 = ym[t];
          T2 yv = y0;
          yv.x += wm * (E.x.x - yv.x);
          yv.y += wm * (E.x.y - yv.y);
          // the smoothed iterate of a lane not running stays as it is (a select per component:
          // a select of the struct went through scratch memory)
          T2 yo;
          yo.x = run ? yv.x : y0.x;
          yo.y = run ? yv.y : y0.y;
          ym[t] = yo;
        }
        const T ur = E.r.x - smv.x, ui = E.r.y - smv.y;
        smv.x += wm * ur;
        smv.y += wm * ui;
        st_nt(sm + t, smv);
        v[2] += double(smv.x) * double(smv.x) + double(smv.y) * double(smv.y);
      } else if (mode == MRS_COPY) {
        if (!LAZY && ym) ym[t] = E.x;
        smv = E.r;
        st_nt(sm + t, smv);
        v[2] += double(E.r.x) * double(E.r.x) + double(E.r.y) * double(E.r.y);
      }
      if (!step2) return;
      if (!LAZY) {
        const T2 pn = caxpy(E.s, E.p, br, bi);
        // x of a lane not running is stored exactly as loaded (dead inputs may hold anything)
        const T2 xn = run ? caxpy(E.x, pn, ar, ai) : E.x;
        st_nt(p + t, pn);
        st_nt(xout + t, xn);
      }
      const T2 rn = caxpy(E.r, E.a, -ar, -ai);
      st_nt(s_vec + t, caxpy(E.s, E.w, -ar, -ai));
      st_nt(r + t, rn);
      const T ur = rn.x - smv.x, ui = rn.y - smv.y;
      v[0] += double(smv.x) * double(ur) + double(smv.y) * double(ui);
      v[1] += double(ur) * double(ur) + double(ui) * double(ui);
    };
    if constexpr (RECW) {
      // one element per thread and step: the help and D_A^-1 of a second element in flight
      // took the kernel to 106 VGPRs (4 waves per SIMD); with one, 76 VGPRs and 6 waves, and
      // the pass runs 6% faster. The partial sums take the elements in the same order.
      for (long long tile = tw.first; tile < tw.last; tile += tw.step) {
        const long long t0 = tile * kBlock + threadIdx.x;
        if (t0 < N) {
          Elem e0;
          load(t0, e0);
          update(t0, e0);
        }
      }
      return;
    }
    for (long long tile = tw.first; tile < tw.last; tile += 2 * tw.step) {
      const long long t0 = tile * kBlock + threadIdx.x;
      const long long t1 = (tile + tw.step) * kBlock + threadIdx.x;
      const bool ok0 = t0 < N;
      const bool ok1 = tile + tw.step < tw.last && t1 < N;
      Elem e0, e1;
      if (ok0) load(t0, e0);
      if (ok1) load(t1, e1);
      if (ok0) update(t0, e0);
      if (ok1) update(t1, e1);
    }
  };
  if (*nrun == SP)
    pass(std::true_type{});
  else
    pass(std::false_type{});
  block_partials<SP, 3>(v, partials);
}


// The x/p recurrence of NP (8 or 16) COCR iterations k-NP+1..k in one pass
// (pcocr_rci.cpp:140-142,164: p_j = s_{j-1} + beta_j p_{j-1}, x_j = x_{j-1} + alpha_j p_j), for
// the FP64 solver whose update pass leaves x and p out (k_xr_fused<.., true>). Nothing else in
// an iteration reads x or p, so they can lag; the update pass keeps s_{k-NP}..s_{k-1} in a ring
// of 2 NP buffers (ring.s[i] below is s_{k-NP+i}) and the scalar kernels record each
// iteration's alpha and beta (SysScalars::hab, position (j - 1) % 32). The arithmetic is that
// of the eager update pass, in the same order, so x and p carry the same bits.
//
// Runs after s_mrs_stop of iteration k (k = nrun[1], a multiple of NP), when every stop up to
// iteration k - 1 is known: a system that stopped at iteration `last` takes the steps j <= last
// only, so x then holds its answer, x_last. A system still running takes all NP; should it
// stop at iteration k, x_k is its answer. Lanes with nothing to apply store x as loaded (their
// dead p and s may hold anything). x and p are updated in place: each element is read and
// written by one thread. It runs even when no system is running any more (at most twice per
// solve), since the last stops of a batch complete their x here.
struct XpRing {
  const double2* s[kXpMax];
};

template <int SP, int NP>
__global__ __launch_bounds__(kBlock) void k_flush_xp(
    double2* __restrict__ x, double2* __restrict__ p, XpRing ring,
    const SysScalars* __restrict__ sc, const int* __restrict__ nrun, int nb) {
  static_assert(NP == 8 || NP == 16, "x/p periods of 8 or 16 iterations");
  const long long N = (long long)nb * SP;
  const long long ntiles = (N + kBlock - 1) / kBlock;
  TileWalk tw(ntiles);
  const int s = int(threadIdx.x % SP);  // kBlock is a multiple of SP: one system per thread
  const int k = nrun[1];
  const bool run = sc[s].status == ST_RUNNING;
  const int last = sc[s].last;
  // alpha and beta of the NP iterations per system in LDS (in registers, eight of them took
  // the kernel to 172 VGPRs), the steps a lane applies as a bit mask
  __shared__ double2 hab_l[NP][SP][2];
  for (int q = threadIdx.x; q < NP * SP; q += kBlock) {
    const int i = q / SP, sy = q % SP;
    const double* hq = sc[sy].hab[(k - NP + i) & (kRingMax - 1)];  // iteration k - NP + 1 + i
    hab_l[i][sy][0] = make_double2(hq[0], hq[1]);
    hab_l[i][sy][1] = make_double2(hq[2], hq[3]);
  }
  unsigned use = 0;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const int j = k - (NP - 1) + i;
    if (j >= 1 && (run || j <= last)) use |= 1u << i;
  }
  __syncthreads();
  // one element per thread and step (x, p and eight s in flight)
  for (long long tile = tw.first; tile < tw.last; tile += tw.step) {
    const long long t = tile * kBlock + threadIdx.x;
    if (t >= N) continue;
    double2 sv[NP];
    double2 xv = ld_nt(x + t), pv = ld_nt(p + t);
#pragma unroll
    for (int i = 0; i < NP; ++i) sv[i] = ld_nt(ring.s[i] + t);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double2 al = hab_l[i][s][0], be = hab_l[i][s][1];
      const double2 pn = caxpy(sv[i], pv, be.x, be.y);
      const double2 xn = caxpy(xv, pn, al.x, al.y);
      // selects, not branches: every lane of a wave keeps issuing the same memory instructions
      const bool u = (use >> i) & 1u;
      pv.x = u ? pn.x : pv.x;
      pv.y = u ? pn.y : pv.y;
      xv.x = u ? xn.x : xv.x;
      xv.y = u ? xn.y : xv.y;
    }
    st_nt(x + t, xv);
    st_nt(p + t, pv);
  }
}

// ---- batch compaction: narrowing the batch to the systems still running -------------------

// dst[row][j] = src[row][map[j]] for the sp_to slots of a narrower batch (map[j] < 0: zero)
template <class T2>
__global__ void k_repack(const T2* __restrict__ src, int sp_from, T2* __restrict__ dst,
                         int sp_to, const int* __restrict__ map, long long rows) {
  const long long N = rows * sp_to;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (long long)gridDim.x * blockDim.x) {
    const long long row = t / sp_to;
    const int m = map[t % sp_to];
    T2 v;
    v.x = 0;
    v.y = 0;
    if (m >= 0) v = src[row * sp_from + m];
    dst[t] = v;
  }
}

// x64[row][sel[j]] = x of slot j's last iteration (buffer last % 2), for the slots with
// sel[j] >= 0 (the systems that finished, in the caller's batch order)
template <int SP>
__global__ void k_scatter_x(double2* __restrict__ x64, int sp0, const double2* __restrict__ x0,
                            const double2* __restrict__ x1, const SysScalars* __restrict__ sc,
                            const int* __restrict__ sel, int nb) {
  const long long N = (long long)nb * SP;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (long long)gridDim.x * blockDim.x) {
    const int j = int(t % SP);
    const int o = sel[j];
    if (o < 0) continue;
    const long long row = t / SP;
    x64[row * sp0 + o] = (sc[j].last & 1) ? x1[t] : x0[t];
  }
}

// ---- scalar kernels (one block) ----------------------------------------------------------

// after the initial residual: r_old = ||b||, X0 test, MRS residual, beta = 0
template <int SP>
__global__ __launch_bounds__(kScalarThreads) void s_prologue(const double* part, int nblk,
                                                             SysScalars* sc, int* nrun) {
  __shared__ double tot[SP * 2];
  reduce_partials<SP * 2>(part, nblk, tot);
  const int s = threadIdx.x;
  if (s < SP && sc[s].status == ST_RUNNING) {
    const double bb = tot[s * 2], rr = tot[s * 2 + 1];
    sc[s].r_old = bb > 0.0 ? sqrt(bb) : 0.0;
    sc[s].b_norm = sc[s].r_old;
    const double rn = rr > 0.0 ? sqrt(rr) : 0.0;
    sc[s].res = rn;  // MRS init: residual = ||r||
    sc[s].rel = 1.0;
    sc[s].best_rel = 1.0;
    sc[s].last_improve = 0;
    sc[s].be_re = sc[s].be_im = 0.0;
    sc[s].iter = 0;
    sc[s].mode = MRS_NONE;  // no MRS step pending before iteration 1
    sc[s].halt = 0;
    sc[s].last = 0;
    if (rn < 1e-30) sc[s].status = MTCOCR_X0_NULL;
  }
  __syncthreads();
  if (threadIdx.x == 0) count_running(sc, SP, nrun);
}

// after k_pc_edge_iter: aw, zero test, alpha = zs / aw (pcocr_rci.cpp:288-299)
template <int SP>
__global__ __launch_bounds__(kScalarThreads) void s_alpha(const double* part, int nblk,
                                                          SysScalars* sc, int* nrun) {
  __shared__ double tot[SP * 2];
  reduce_partials<SP * 2>(part, nblk, tot);
  const int s = threadIdx.x;
  if (s < SP && sc[s].status == ST_RUNNING && sc[s].halt == 0) {
    const double awr = tot[s * 2], awi = tot[s * 2 + 1];
    sc[s].aw_re = awr;
    sc[s].aw_im = awi;
    if (fabs(awr) < 1e-40 && fabs(awi) < 1e-40) {
      // takes effect after the previous iteration's deferred stop test (s_mrs_stop)
      sc[s].halt = MTCOCR_DIV_BY_ZERO;
    } else {
      double qr, qi;
      cdiv(sc[s].zs_re, sc[s].zs_im, awr, awi, qr, qi);
      sc[s].al_re = qr;
      sc[s].al_im = qi;
      double* hq = sc[s].hab[(sc[s].iter - 1) & (kRingMax - 1)];
      hq[0] = qr;
      hq[1] = qi;
      hq[2] = sc[s].be_re;
      hq[3] = sc[s].be_im;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) count_running(sc, SP, nrun);
}

// After k_xr_fused of iteration k: first the stop test of iteration k - 1, whose MRS step the
// fused pass has just completed (rci.cpp:66-77 on the smoothed residual ||sm||), then a halt
// decided during iteration k (division by zero, iteration limit), then the MRS weight of
// iteration k, w = -(sm,u)/(u,u) clamped (MRS.cpp:113-121), applied by the next fused pass.
template <int SP>
__global__ __launch_bounds__(kScalarThreads) void s_mrs_stop(const double* part, int nblk,
                                                             SysScalars* sc, int* nrun) {
  __shared__ double tot[SP * 3];
  reduce_partials<SP * 3>(part, nblk, tot);
  const int s = threadIdx.x;
  if (s < SP && sc[s].status == ST_RUNNING) {
    const double su = tot[s * 3], uu = tot[s * 3 + 1], ss = tot[s * 3 + 2];
    const int k = sc[s].iter;  // this iteration; the pending MRS step belongs to k - 1
    bool stopped = false;
    const int prev = sc[s].mode;
    if (prev != MRS_NONE) {
      if (prev == MRS_BLEND) sc[s].res = sqrt(ss);
      else if (prev == MRS_COPY) sc[s].res = ss > 0.0 ? sqrt(ss) : 0.0;
      sc[s].rel = sc[s].res / sc[s].r_old;
      if (sc[s].rel < sc[s].eps) {
        sc[s].status = MTCOCR_OK;
        sc[s].iter = k - 1;
        sc[s].last = k - 1;
        stopped = true;
      } else if (sc[s].stall_window > 0) {
        // inner solves of the refinement stop when the smoothed residual has not halved within
        // stall_window iterations (FP32 has reached its attainable accuracy)
        if (sc[s].rel < 0.5 * sc[s].best_rel) {
          sc[s].best_rel = sc[s].rel;
          sc[s].last_improve = k - 1;
        } else if (k - 1 - sc[s].last_improve > sc[s].stall_window) {
          sc[s].status = MTCOCR_MAXITER;
          sc[s].last = k - 1;
          stopped = true;
        }
      }
    }
    if (!stopped) {
      if (sc[s].halt) {
        sc[s].status = sc[s].halt;  // x of iteration k - 1 is the answer
        sc[s].last = k - 1;
      } else {
        double w = (fabs(uu) < 1e-100) ? 0.0 : -su / uu;
        int mode;
        if (w < 0) {
          w = 0;
          mode = MRS_SKIP;
        } else if (w > 1) {
          w = 1;
          mode = MRS_COPY;
        } else {
          mode = MRS_BLEND;
        }
        sc[s].w = w;
        sc[s].mode = mode;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) count_running(sc, SP, nrun);
}

// after k_spmv (z = A s): zs, beta, next iteration (pcocr_rci.cpp:263-270,323-337)
template <int SP, bool FIRST>
__global__ __launch_bounds__(kScalarThreads) void s_beta(const double* part, int nblk,
                                                         SysScalars* sc, int* nrun, int maxiter) {
  __shared__ double tot[SP * 2];
  reduce_partials<SP * 2>(part, nblk, tot);
  const int s = threadIdx.x;
  if (s < SP && sc[s].status == ST_RUNNING) {
    const double zr = tot[s * 2], zi = tot[s * 2 + 1];
    if (FIRST) {
      sc[s].zs_re = zr;
      sc[s].zs_im = zi;
      sc[s].be_re = sc[s].be_im = 0.0;
    } else {
      sc[s].zs1_re = sc[s].zs_re;
      sc[s].zs1_im = sc[s].zs_im;
      sc[s].zs_re = zr;
      sc[s].zs_im = zi;
      // the reference tests aw here, not zs (pcocr_rci.cpp:328)
      if (fabs(sc[s].aw_re) < 1e-40 && fabs(sc[s].aw_im) < 1e-40) {
        if (!sc[s].halt) sc[s].halt = MTCOCR_DIV_BY_ZERO;
      } else {
        double qr, qi;
        cdiv(zr, zi, sc[s].zs1_re, sc[s].zs1_im, qr, qi);
        sc[s].be_re = qr;
        sc[s].be_im = qi;
      }
    }
    if (sc[s].status == ST_RUNNING && !sc[s].halt) {
      sc[s].iter += 1;
      // the stop test of the iteration just finished is still pending (s_mrs_stop of the next
      // iteration decides it first)
      if (sc[s].iter > maxiter) sc[s].halt = MTCOCR_MAXITER;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    count_running(sc, SP, nrun);
    // nrun[1]: the batch's iteration number (that of every running system), which
    // k_flush_xp compares with the iteration of a system that has stopped
    nrun[1] = FIRST ? 1 : nrun[1] + 1;
  }
}

// First stage of the partial-sum reduction for wide grids (more than kFoldAbove partial
// doubles): workgroup g folds partial rows
// [g*fold, (g+1)*fold) into out[g*P + q], in a fixed order, so that the one-workgroup scalar
// kernels read at most about kFoldRows rows (one CU cannot stream megabytes of partials
// quickly). fold = max(kFold, ceil(nblk / kFoldRows)) depends on the grid only, so the sums
// are deterministic like the single-stage sum, with a different grouping.
constexpr int kFold = 64;
constexpr int kFoldRows = 256;
constexpr long long kFoldAbove = 16384;  // partial doubles (rows x P) reduced in one stage

inline int fold_factor(int nblk) { return std::max(kFold, (nblk + kFoldRows - 1) / kFoldRows); }

template <int P>
__global__ __launch_bounds__(256) void k_fold(const double* __restrict__ part, int nblk, int fold,
                                              double* __restrict__ out) {
  constexpr int R = 256 / P > 0 ? 256 / P : 1;  // partial rows handled side by side
  constexpr int U = 16;                         // loads of a thread in flight together
  __shared__ double sh[256];
  const int g = blockIdx.x;
  const int q = threadIdx.x % P, j = threadIdx.x / P;
  const int b0 = g * fold, b1 = min(b0 + fold, nblk);
  double acc = 0.0;
  if (threadIdx.x < R * P) {
    for (int bc = b0 + j; bc < b1; bc += U * R) {
      double v[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int b = bc + k * R;
        v[k] = b < b1 ? part[(size_t)b * P + q] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < U; ++k) acc += v[k];
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < P) {
    double t = 0.0;
    for (int jj = 0; jj < R; ++jj) t += sh[jj * P + threadIdx.x];
    out[(size_t)g * P + threadIdx.x] = t;
  }
}

// the partial-sum reduction of a scalar kernel alone (the autotuner times kernel + reduction)
template <int P>
__global__ __launch_bounds__(kScalarThreads) void s_reduce_probe(const double* part, int nblk,
                                                                 double* out) {
  __shared__ double tot[P];
  reduce_partials<P>(part, nblk, tot);
  if (threadIdx.x < P) out[threadIdx.x] = tot[threadIdx.x];
}

// state of one running system (SP = 1) for the allocation-free entry point: its scalars, the
// running count and omega, written on the device (no host copy, so that the launch sequence
// can be captured)
__global__ void k_noalloc_init(SysScalars* sc, int* nrun, double* om, double omega) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    SysScalars z = SysScalars();
    z.status = ST_RUNNING;
    z.outer_status = ST_RUNNING;
    z.rel = 1.0;
    *sc = z;
    *nrun = 1;
    *om = omega;
  }
}

// ---- mixed precision: FP64 outer refinement ------------------------------------------------

// after the FP64 true residual: relative residual, convergence / stall, inner set-up
template <int SP>
__global__ __launch_bounds__(kScalarThreads) void s_outer(const double* part, int nblk,
                                                          SysScalars* sc, int* nrun_outer,
                                                          int* nrun_inner, double eps,
                                                          double inner_floor, int first,
                                                          int max_outer) {
  __shared__ double tot[SP * 2];
  reduce_partials<SP * 2>(part, nblk, tot);
  const int s = threadIdx.x;
  if (s < SP) {
    sc[s].rollback = 0;
    if (sc[s].outer_status == ST_RUNNING) {
      const double bb = tot[s * 2], rr = tot[s * 2 + 1];
      const double rn = rr > 0.0 ? sqrt(rr) : 0.0;
      if (first) {
        sc[s].b_norm = bb > 0.0 ? sqrt(bb) : 0.0;
        sc[s].prev_rel = 1e300;
        sc[s].outer = 0;
        sc[s].total_iter = 0;
      }
      const double bn = sc[s].b_norm;
      if (rn < 1e-30) {
        sc[s].outer_status = first ? MTCOCR_X0_NULL : MTCOCR_OK;
        sc[s].rel = (bn > 0.0) ? rn / bn : 0.0;
      } else {
        const double rel = rn / bn;
        sc[s].rel = rel;
        if (rel < eps) {
          sc[s].outer_status = MTCOCR_OK;
        } else if (!first && rel > 0.1 * sc[s].prev_rel) {
          // a refinement pass that gains less than 10x hands the system to the FP64 solver;
          // one that made the residual worse is undone first (k_rollback)
          sc[s].outer_status = MTCOCR_ERR_NOTCONV;
          sc[s].rollback = rel > sc[s].prev_rel ? 1 : 0;
          if (sc[s].rollback) sc[s].rel = sc[s].prev_rel;
        } else if (sc[s].outer >= max_outer) {
          sc[s].outer_status = MTCOCR_ERR_NOTCONV;
        } else {
          sc[s].prev_rel = rel;
          sc[s].scale = 1.0 / rn;
          // inner tolerance: what is still missing, with a safety factor, floored at the
          // attainable FP32 level
          double tau = 0.1 * eps / rel;
          if (tau < inner_floor) tau = inner_floor;
          if (tau > 0.1) tau = 0.1;
          sc[s].eps = tau;
          sc[s].outer += 1;
        }
      }
    }
    sc[s].status = (sc[s].outer_status == ST_RUNNING) ? ST_RUNNING : ST_INACTIVE;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int q = 0; q < SP; ++q) c += (sc[q].outer_status == ST_RUNNING);
    *nrun_outer = c;
    *nrun_inner = c;
  }
}

// inner right-hand side b32 = r64 * scale, x32 = 0
template <int SP>
__global__ void k_to_inner(const double2* r64, float2* b32, float2* x32, const SysScalars* sc,
                           int nb) {
  const long long N = (long long)nb * SP;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (long long)gridDim.x * blockDim.x) {
    const int s = int(t % SP);
    if (sc[s].outer_status != ST_RUNNING) continue;
    const double sc_ = sc[s].scale;
    float2 o;
    o.x = float(r64[t].x * sc_);
    o.y = float(r64[t].y * sc_);
    b32[t] = o;
    x32[t] = make_float2(0.f, 0.f);
  }
}

// x64 += y32 / scale (the MRS-smoothed inner iterate)
template <int SP>
__global__ void k_from_inner(double2* x64, const float2* y32, SysScalars* sc, int nb) {
  const long long N = (long long)nb * SP;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (long long)gridDim.x * blockDim.x) {
    const int s = int(t % SP);
    if (sc[s].outer_status != ST_RUNNING) continue;
    const double inv = 1.0 / sc[s].scale;
    x64[t].x += double(y32[t].x) * inv;
    x64[t].y += double(y32[t].y) * inv;
  }
}

// undo the last correction x64 += y32 / scale of the systems whose pass made things worse
template <int SP>
__global__ void k_rollback(double2* x64, const float2* y32, const SysScalars* sc, int nb) {
  const long long N = (long long)nb * SP;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (long long)gridDim.x * blockDim.x) {
    const int s = int(t % SP);
    if (!sc[s].rollback) continue;
    const double inv = 1.0 / sc[s].scale;
    x64[t].x -= double(y32[t].x) * inv;
    x64[t].y -= double(y32[t].y) * inv;
  }
}

template <int SP>
__global__ void s_inner_done(SysScalars* sc) {
  const int s = threadIdx.x;
  if (s < SP && sc[s].outer_status == ST_RUNNING) sc[s].total_iter += sc[s].iter;
}

// ------------------------------------------------------------------------------------------
// device memory

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace mtcocr

using namespace mtcocr;

// ------------------------------------------------------------------------------------------
// the handle

struct mtcocr_handle {
  int device = 0;
  int n_cu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int* h_flags = nullptr;  // pinned: [0] running count

  // device row order (new -> reference edge index); empty = the reference's order
  std::vector<int> order;
  // the grid order of a box mesh (mtcocr_set_mesh_order with MTCOCR_ORDER_GRID): its node
  // counts, and the stencil coefficients of the matrix when every row fits the stencil
  bool grid_order = false;
  GridDims grid;
  // k_spmv_stl: tile columns per band of its task order (0: one band), per precision and width
  int stl_band[2][7] = {};
  DevBuf stc;
  bool stc_ok = false;
  // the table is above 4 GiB (each direction's box under it): only the LDS-staged kernel, which
  // addresses it box by box, reads it
  bool stc_big = false;

  // matrix (shared index arrays, FP64 and FP32 values)
  bool have_matrix = false;
  bool ks_form = false;
  int nb = 0, nslices = 0;
  int64_t nnz_c = 0, nnz_r = 0;
  DevBuf offc, nc, offr, nr, colc, colr;
  DevBuf ac64, bc64, ar64, da64, db64;
  DevBuf rp, rb, pcol, pa, pb;  // row pack of the off-diagonal entries
  DevBuf hoff, halo, lcol, ldiag;  // 16-row blocks of the row pack and their halos
  int nblk = 0, halo_max = 0;      // halo_max 0: no row blocks
  // gradient (g_nb: the edge count it was built for; the solver runs only when it equals nb)
  bool have_g = false;
  int m = 0, wg = 0, wgt = 0, g_nb = 0;
  DevBuf gcol, gtcol, gval64, gtval64, gval32, gtval32;
  // G is the whole box's gradient in the grid order (gradient_on_grid): batches of 8 or 16 then
  // run the fused preconditioner kernel k_pc_grid, with its node levels per chunk per precision
  bool pg_ok = false;
  int pg_kc[2] = {8, 8};

  // batch state
  int nsys = 0, sp = 0;
  DevBuf omega, scal, nrun, nrun2, part, part2;
  DevBuf b64, x64, r64;  // FP64 rhs / solution / outer residual
  // COCR work vectors (type depends on the precision of the last solve)
  DevBuf vr, vp, vs, vz, va, vw, vym, vsm, vhelp, vdn, vx, vx2, vb;
  DevBuf va2;  // the second buffer of a, alternating with va, for k_pc_grid
  DevBuf vsr[kRingMax - 1];  // s buffers 1..31 of the FP64 solver's ring (buffer 0 is vs)
  // the FP64 solver's x/p period: 0 chooses 8 or 16 by free device memory at the first FP64
  // solve of a batch shape (mtcocr_set_xp_period), xp_period the one in use
  int xp_request = 0, xp_period = 8;
  // the same vectors for a narrowed batch (swapped with the above by compaction)
  DevBuf cvr, cvp, cvs, cvz, cva, cvw, cvym, cvsm, cvx, cvx2;
  DevBuf cmap;  // slot maps of compaction / scatter
  // the adjoint's gradient contraction (mtcocr_sigma_gradient): cell edges, node coordinates,
  // coefficients, the two vector batches and the per-cell result
  DevBuf sg_ce, sg_xyz, sg_coef, sg_u, sg_v, sg_out;
  int work_prec = -1, work_sp = 0;
  hipGraphExec_t graph = nullptr;
  int graph_prec = -1, graph_sp = 0, graph_chunk = 0, graph_maxiter = -1, graph_pg = -1;
  // launch shapes (workgroup caps of the grid-stride kernels) per vector precision (FP64,
  // FP32) and batch width SP = 1..64 (index log2 SP), set by mtcocr_set_launch / autotune
  int cap_spmv[2][7], cap_vec[2][7], cap_node[2][7];
  // the fused update pass's own cap: set with cap_vec by set_launch, tuned apart by autotune
  int cap_upd[2][7];
  // SpMV layout per precision and batch width: 0 sliced ELL (k_spmv), 1 row pack (k_spmv_rg,
  // batch widths 4..64), 2 halo-staged row blocks (k_spmv_blk, batch widths 4..16; where it
  // does not apply, the row pack runs), 3 structured stencil (k_spmv_stc, batch widths 4..16 on
  // a matrix held in the grid order; else the row blocks)
  int spmv_layout[2][7];
  int stc_kc[2][7];  // the stencil kernel's sweep length (rows of k per wave task)
  static int sp_index(int sp) {
    int k = 0;
    while ((1 << k) < sp && k < 6) ++k;
    return k;
  }
  void set_caps(int cs, int cv, int cn, int layout) {
    for (int p = 0; p < 2; ++p)
      for (int k = 0; k < 7; ++k) {
        if (cs) cap_spmv[p][k] = cs;
        if (cv) cap_vec[p][k] = cap_upd[p][k] = cv;
        if (cn) cap_node[p][k] = cn;
        if (layout >= 0) spmv_layout[p][k] = (k < 2) ? 0 : (k > 4 && layout > 1) ? 1 : layout;
      }
  }
  mtcocr_handle() {
    // the LDS-staged stencil for 16 systems where the matrix has it, else row blocks for 4..16
    // systems where their halos fit, else the row pack (set_caps clamps the layout for 32 and
    // 64); streaming kernels at about one tile per workgroup (the grid is clamped to the tiles)
    set_caps(8192, 65536, 2048, 4);
    for (auto& row : cap_upd)
      for (int& c : row) c = kMaxBlocks;  // the update pass: one tile per workgroup
    for (auto& row : stc_kc)
      for (int& c : row) c = kWholeColumn;  // stencil sweeps over whole columns of the mesh
  }

  int64_t device_bytes() const {
    const DevBuf* all[] = {&offc, &nc, &offr, &nr, &colc, &colr, &ac64, &bc64, &ar64, &da64,
                           &db64, &rp, &rb, &pcol, &pa, &pb, &hoff, &halo, &lcol, &ldiag, &stc, &gcol, &gtcol, &gval64, &gtval64,
                           &gval32, &gtval32, &omega, &scal, &nrun, &nrun2, &part, &part2, &b64, &x64,
                           &r64, &vr, &vp, &vs, &vz, &va, &vw, &vym, &vsm, &vhelp, &vdn, &vx2, &cvr, &cvp, &cvs, &cvz, &cva, &cvw, &cvym,
                           &cvsm, &cvx, &cvx2, &cmap,
                           &vx, &vb, &sg_ce, &sg_xyz, &sg_coef, &sg_u, &sg_v, &sg_out, &va2};
    int64_t t = 0;
    for (auto* b : all) t += (int64_t)b->bytes;
    for (const DevBuf& b : vsr) t += (int64_t)b.bytes;
    return t;
  }
};

// Device scratch of the allocation-free entry point (mtcocr_apply_noalloc): the per-system
// scalars, the SpMV's partial sums and the preconditioner's node vectors live here instead of in
// buffers the handle allocates. Each device holds one copy of the array for the life of the
// process; host code uses the per-device address from hipGetSymbolAddress (scratch_ptr), never
// the array's name, which in host code is only a stand-in.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local char g_err[512];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_rc[kMaxDevices];
bool g_scratch_tried[kMaxDevices];

// Address of this device's scratch array, looked up once per device with the device current
// (the handle's set-up calls it); null when the lookup failed.
void* scratch_ptr(int device, hipError_t* rc_out) {
  if (device < 0 || device >= kMaxDevices) {
    if (rc_out) *rc_out = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_tried[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_rc[device] = e;
    g_scratch_ptr[device] = (e == hipSuccess) ? p : nullptr;
    g_scratch_tried[device] = true;
  }
  if (rc_out) *rc_out = g_scratch_rc[device];
  return g_scratch_ptr[device];
}

#define HIPC(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      std::snprintf(g_err, sizeof(g_err), "%s failed: %s", #expr, hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? MTCOCR_ERR_ALLOC : MTCOCR_ERR_HIP;               \
    }                                                                                     \
  } while (0)

void drop_graph(mtcocr_handle* h) {
  if (h->graph) (void)hipGraphExecDestroy(h->graph);
  h->graph = nullptr;
}

template <class T>
int upload(mtcocr_handle* h, DevBuf& d, const std::vector<T>& v, hipStream_t st) {
  drop_graph(h);
  HIPC(d.alloc(v.size() * sizeof(T)));
  if (!v.empty()) HIPC(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return 0;
}

int upload_cast_float(mtcocr_handle* h, DevBuf& d, const std::vector<double>& v, hipStream_t st) {
  drop_graph(h);
  std::vector<float> f(v.size());
  for (size_t i = 0; i < v.size(); ++i) f[i] = float(v[i]);
  HIPC(d.alloc(f.size() * sizeof(float)));
  if (!f.empty()) HIPC(hipMemcpyAsync(d.p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, st));
  HIPC(hipStreamSynchronize(st));
  return 0;
}

int pad_sp(int nsys) {
  int sp = 1;
  while (sp < nsys) sp <<= 1;
  return sp;
}

template <class T>
EllView<T> ell_view(mtcocr_handle* h) {
  static_assert(std::is_same<T, double>::value, "the matrix is held in FP64 only");
  EllView<T> v;
  v.offc = h->offc.as<int>();
  v.nc = h->nc.as<int>();
  v.offr = h->offr.as<int>();
  v.nr = h->nr.as<int>();
  v.colc = h->colc.as<int>();
  v.colr = h->colr.as<int>();
  v.ac = h->ac64.as<double>();
  v.bc = h->bc64.as<double>();
  v.ar = h->ar64.as<double>();
  v.nb = h->nb;
  return v;
}

RowView row_view(mtcocr_handle* h) {
  RowView v;
  v.rp = h->rp.as<int>();
  v.rb = h->rb.as<int>();
  v.col = h->pcol.as<int>();
  v.a = h->pa.as<double>();
  v.b = h->pb.as<double>();
  v.da = h->da64.as<double>();
  v.db = h->db64.as<double>();
  v.nb = h->nb;
  return v;
}

BlkView blk_view(mtcocr_handle* h) {
  BlkView v;
  v.rp = h->rp.as<int>();
  v.rb = h->rb.as<int>();
  v.hoff = h->hoff.as<int>();
  v.halo = h->halo.as<int>();
  v.lcol = h->lcol.as<unsigned short>();
  v.ldiag = h->ldiag.as<unsigned short>();
  v.a = h->pa.as<double>();
  v.b = h->pb.as<double>();
  v.da = h->da64.as<double>();
  v.db = h->db64.as<double>();
  v.nb = h->nb;
  v.nblk = h->nblk;
  v.nhalo = (long long)(h->halo.bytes / sizeof(int));
  v.nnz = (long long)(h->pa.bytes / sizeof(double));
  v.nnzb = (long long)(h->pb.bytes / sizeof(double));
  return v;
}

// LDS bytes of one k_spmv_blk workgroup (16/SP blocks of halo_max x SP values), or 0 when
// the matrix has no row blocks or one has more than kHaloStage columns
template <class TV>
int blk_lds_bytes(const mtcocr_handle* h) {
  if (h->halo_max <= 0 || h->halo_max > kHaloStage) return 0;
  return kRowBlock * h->halo_max * int(sizeof(typename Vec2<TV>::type));
}

// view of the stencil layout for batch width SP (RPW = 64 / SP rows of a wave) and sweep kc
template <int SP>
StcView stc_view(const mtcocr_handle* h, int kc) {
  StcView v;
  v.coef = h->stc.as<double>();
  v.nx = h->grid.nx;
  v.ny = h->grid.ny;
  v.nz = h->grid.nz;
  v.kc = std::max(1, kc);
  constexpr int RPW = 64 / SP;
  v.nic = (h->grid.nx + RPW - 1) / RPW;
  const long long nkc = (h->grid.nz + v.kc - 1) / v.kc;
  v.ntasks = 3LL * nkc * h->grid.ny * v.nic;
  v.nb = h->nb;
  return v;
}

// the same for the LDS-staged kernel: tasks are 4 x kStlNW tiles of rows (one per workgroup)
inline StcView stl_view(const mtcocr_handle* h, int kc, int band) {
  StcView v = stc_view<16>(h, kc);
  v.nic = (h->grid.nx + kStlTI - 1) / kStlTI;
  v.njc = (h->grid.ny + kStlNW - 1) / kStlNW;
  v.nbw = band > 0 ? std::min(band, v.nic) : v.nic;
  const long long nkc = (h->grid.nz + v.kc - 1) / v.kc;
  v.ntasks = 3LL * nkc * v.njc * v.nic;
  return v;
}

template <class TV, int SP>
int stc_kc_of(const mtcocr_handle* h) {
  return h->stc_kc[std::is_same<TV, float>::value ? 1 : 0][mtcocr_handle::sp_index(SP)];
}

template <class TV, int SP>
int stl_band_of(const mtcocr_handle* h) {
  return h->stl_band[std::is_same<TV, float>::value ? 1 : 0][mtcocr_handle::sp_index(SP)];
}

// The row kernels address x with 32-bit byte offsets: batch vectors of nb x SP values of
// `elem` bytes must stay under 4 GiB (else the sliced ELL runs)
inline bool rows_fit(const mtcocr_handle* h, int sp, int elem) {
  return (long long)h->nb * sp * elem <= (long long)kOob;
}

// SpMV launch in the layout chosen for this batch width (vector type TV, which the outer FP64
// residual of the mixed solver uses with an FP32 engine).
template <class TV, int SP, int MODE>
void launch_spmv(mtcocr_handle* h, int layout, int grid, hipStream_t st, const double* omega,
                 const typename Vec2<TV>::type* x, typename Vec2<TV>::type* y,
                 const typename Vec2<TV>::type* b, typename Vec2<TV>::type* ym,
                 typename Vec2<TV>::type* sm, const SysScalars* sc, const int* nrun,
                 double* part) {
  if (!rows_fit(h, SP, int(sizeof(typename Vec2<TV>::type)))) layout = 0;
  if (layout == 3 && h->stc_big) layout = 4;  // the wave-task kernel cannot address the table
  if constexpr (SP >= 4 && SP <= 16) {
    if constexpr (SP == 8) {
      if (layout == 4 && h->stc_ok) {
        k_spmv_stl<TV, SP, MODE><<<grid, kBlock, 0, st>>>(
            stl_view(h, stc_kc_of<TV, SP>(h), stl_band_of<TV, SP>(h)), omega, x, y, b, ym, sm,
            sc, nrun, part);
        return;
      }
    }
    if constexpr (SP == 16) {
      if (layout == 3 && h->stc_ok && !h->stc_big) {
        k_spmv_stc<TV, SP, MODE><<<grid, kBlock, 0, st>>>(stc_view<SP>(h, stc_kc_of<TV, SP>(h)),
                                                          omega, x, y, b, ym, sm, sc, nrun, part);
        return;
      }
      if (layout == 4 && h->stc_ok) {
        k_spmv_stl<TV, SP, MODE><<<grid, kBlock, 0, st>>>(
            stl_view(h, stc_kc_of<TV, SP>(h), stl_band_of<TV, SP>(h)), omega, x, y, b, ym, sm,
            sc, nrun, part);
        return;
      }
    }
    if (layout >= 3) layout = 2;  // the stencil layouts serve 16 (and, LDS-staged, 8) systems
    const int lds = blk_lds_bytes<TV>(h);
    if (layout == 2 && lds > 0) {
      k_spmv_blk<TV, SP, MODE><<<grid, kBlock, lds, st>>>(blk_view(h), h->halo_max, omega, x, y,
                                                          b, ym, sm, sc, nrun, part);
      return;
    }
  }
  if constexpr (SP >= 4) {
    if (layout >= 1) {
      k_spmv_rg<TV, SP, MODE><<<grid, kBlock, 0, st>>>(row_view(h), omega, x, y, b, ym, sm, sc,
                                                       nrun, part);
      return;
    }
  }
  k_spmv<TV, SP, MODE><<<grid, kBlock, 0, st>>>(ell_view<double>(h), omega, x, y, b, ym, sm, sc,
                                                nrun, part);
}

template <class T>
GView<T> g_view(mtcocr_handle* h) {
  GView<T> g;
  g.gcol = h->gcol.as<int>();
  g.gtcol = h->gtcol.as<int>();
  if constexpr (std::is_same<T, double>::value) {
    g.gval = h->gval64.as<double>();
    g.gtval = h->gtval64.as<double>();
  } else {
    g.gval = h->gval32.as<float>();
    g.gtval = h->gtval32.as<float>();
  }
  g.m = h->m;
  g.n = h->nb;
  g.wg = h->wg;
  g.wgt = h->wgt;
  return g;
}

// The fused grid preconditioner applies to batches of 8 or 16 systems on a box mesh in the grid order
// whose G has the box's structure, with both a buffers allocated, byte offsets of the batch
// vectors within 32 bits, and unless MTCOCR_PC_SPLIT=1 asks for the node and edge kernels.
inline bool pg_applies(const mtcocr_handle* h, int sp, size_t es) {
  if (!pg_width(sp) || !h->pg_ok || !h->have_g) return false;
  if ((long long)h->nb * sp * (long long)es >= (long long)kOob) return false;
  if (h->va2.bytes < (size_t)h->nb * sp * es || h->va.bytes < (size_t)h->nb * sp * es) return false;
  const char* e = std::getenv("MTCOCR_PC_SPLIT");
  return !(e && e[0] == '1');
}

inline PgView pg_view(const mtcocr_handle* h, int kc) {
  PgView v;
  v.nx = h->grid.nx;
  v.ny = h->grid.ny;
  v.nz = h->grid.nz;
  v.nti = (v.nx + kPgTI - 1) / kPgTI;
  v.ntj = (v.ny + kPgTJ - 1) / kPgTJ;
  v.kc = std::max(1, std::min(kc, v.nz));
  v.nkc = (v.nz + v.kc - 1) / v.kc;
  v.ntasks = (long long)v.nkc * v.ntj * v.nti;
  v.nb = h->nb;
  return v;
}

// one workgroup per task up to the partials' capacity, a multiple of the XCD count
inline int pg_grid(const PgView& v) {
  long long g = std::min<long long>(v.ntasks, kMaxBlocks);
  g = std::max<long long>(g, kXcd);
  return round_up(int(g), kXcd);
}

// COCR engine over SP systems in storage precision T.
template <class T, int SP>
struct Engine {
  using T2 = typename Vec2<T>::type;
  mtcocr_handle* h;
  EllView<double> A;
  GView<T> G;
  T2 *x, *x2, *b, *r, *p, *s, *z, *a, *a2, *w, *ym, *sm, *help, *dn;
  DiagA da;
  SysScalars* sc;
  int* nrun;
  double *part, *part2;
  const double* omega;
  int nb, m, grid_s, grid_e, grid_n, grid_u, layout;
  // the FP64 solver's ring of s buffers (s_k in sr[k % 16], sr[0] = s): its update pass leaves
  // x and p to k_flush_xp, which reads s_{k-8}..s_{k-1} every eighth iteration
  static constexpr bool kLazy = std::is_same<T, double>::value;
  T2* sr[kRingMax];
  int xp = 8, ring = 16;  // x/p period and ring length of this handle
  // the fused grid preconditioner (k_pc_grid): used, its view and grid
  bool use_pg = false;
  PgView pg;
  int grid_pg = 0;

  explicit Engine(mtcocr_handle* hh) : h(hh) {
    A = ell_view<double>(h);
    G = g_view<T>(h);
    nb = h->nb;
    m = h->m;
    x = h->vx.as<T2>();
    x2 = h->vx2.as<T2>();
    b = h->vb.as<T2>();
    r = h->vr.as<T2>();
    p = h->vp.as<T2>();
    s = h->vs.as<T2>();
    z = h->vz.as<T2>();
    a = h->va.as<T2>();
    w = h->vw.as<T2>();
    ym = h->vym.as<T2>();
    sm = h->vsm.as<T2>();
    help = h->vhelp.as<T2>();
    da.a = h->da64.as<double>();
    da.b = h->db64.as<double>();
    da.omega = h->omega.as<double>();
    dn = h->vdn.as<T2>();
    sc = h->scal.as<SysScalars>();
    nrun = h->nrun.as<int>();
    part = h->part.as<double>();
    part2 = h->part2.as<double>();
    omega = h->omega.as<double>();
    const int k = mtcocr_handle::sp_index(SP), q = std::is_same<T, float>::value ? 1 : 0;
    grid_s = grid_for((long long)nb * SP, h->cap_spmv[q][k]);
    grid_e = grid_for((long long)nb * SP, h->cap_vec[q][k]);
    grid_u = grid_for((long long)nb * SP, h->cap_upd[q][k]);
    grid_n = grid_for((long long)m * SP, h->cap_node[q][k]);
    layout = h->spmv_layout[q][k];
    a2 = h->va2.as<T2>();
    sr[0] = s;
    xp = h->xp_period;
    ring = 2 * xp;
    for (int i = 1; i < kRingMax; ++i) sr[i] = kLazy && i < ring ? h->vsr[i - 1].as<T2>() : s;
    use_pg = pg_applies(h, SP, sizeof(T2));
    if (use_pg) {
      pg = pg_view(h, h->pg_kc[q]);
      grid_pg = pg_grid(pg);
    }
  }

  // what the update pass needs to form w after the fused preconditioner (its box, D_A)
  WRec wrec() const {
    WRec q;
    q.nx = pg.nx;
    q.ny = pg.ny;
    q.nz = pg.nz;
    q.da = da;
    return q;
  }

  // The MRS-smoothed iterate: the FP32 inner solve of the mixed solver returns it as its
  // correction; the FP64 solver returns x unsmoothed, as the reference does (its MRS keeps y
  // but nothing reads it back), so there it is not kept (null: no traffic).
  T2* ym_or_null() const { return std::is_same<T, float>::value ? ym : nullptr; }

  // partials of a grid of nblk workgroups, P values each, as the scalar kernels read them:
  // folded by k_fold into part2 first when the grid is wide
  template <int P>
  const double* partials(hipStream_t st, int nblk, int& nread) {
    if ((long long)nblk * P <= kFoldAbove) {
      nread = nblk;
      return part;
    }
    const int fold = fold_factor(nblk);
    nread = (nblk + fold - 1) / fold;
    k_fold<P><<<nread, 256, 0, st>>>(part, nblk, fold, part2);
    return part2;
  }

  template <int MODE, class TV>
  void spmv(hipStream_t st, const TV* xin, TV* yout, const TV* bin, TV* ymo, TV* smo) {
    using E = decltype(xin->x);
    launch_spmv<E, SP, MODE>(h, layout, grid_s, st, omega, xin, yout, bin, ymo, smo, sc, nrun,
                             part);
  }

  // D_N^-1 per system (D_A^-1 is evaluated inside the edge kernels)
  void setup_diagonals(hipStream_t st) {
    const int g2 = std::min(4096, (int)(((long long)m + 255) / 256) + 1);
    k_diag_n<T, SP><<<g2, 256, 0, st>>>(ell_view<double>(h), g_view<double>(h), omega, dn);
  }

  // stages 1-4 of the request order: r = b - A x, X0 test, s = M^-1 r, p = a = w = 0, z = A s
  void prologue(hipStream_t st, int maxiter) {
    spmv<1>(st, x, r, b, ym_or_null(), sm);
    int nr;
    const double* pp = partials<SP * 2>(st, grid_s, nr);
    s_prologue<SP><<<1, kScalarThreads, 0, st>>>(pp, nr, sc, nrun);
    k_pc_node<T, SP, false><<<grid_n, kBlock, 0, st>>>(G, r, nullptr, dn, help, sc, nrun);
    k_pc_edge_plain<T, SP><<<grid_e, kBlock, 0, st>>>(G, r, da, help, s, sc, nrun);
    const long long nv = (long long)nb * SP;
    const int gf = std::min<long long>(2048, (nv + 255) / 256);
    k_fill<T2><<<gf, 256, 0, st>>>(p, nv);
    k_fill<T2><<<gf, 256, 0, st>>>(a, nv);
    k_fill<T2><<<gf, 256, 0, st>>>(w, nv);
    spmv<0>(st, s, z, (const T2*)nullptr, (T2*)nullptr, (T2*)nullptr);
    pp = partials<SP * 2>(st, grid_s, nr);
    s_beta<SP, true><<<1, kScalarThreads, 0, st>>>(pp, nr, sc, nrun, maxiter);
  }

  // one COCR iteration (stages 5-8) at position c of a graph chunk (c = 0..15 with
  // c = (k - 1) % 16 for iteration k in FP64, c = 0..7 in FP32). `odd` = k is odd: the eager
  // update pass then reads x from buffer 0 and writes buffer 1, and k_pc_grid reads a and
  // writes a2.
  void iteration(hipStream_t st, int maxiter, int c) {
    const bool odd = (c & 1) == 0;
    int nr;
    const double* pp;
    T2* acur = a;  // a of this iteration after the preconditioner pass
    if constexpr (pg_width(SP)) {
      if (use_pg) {
        // a alternates between the two buffers: odd iterations read a and write a2
        acur = odd ? a2 : a;
        if constexpr (kLazy)  // help to the node vector; the update pass forms w from it
          k_pc_grid<T, SP, true><<<grid_pg, kBlock, 0, st>>>(pg, z, odd ? a : a2, acur, help, dn,
                                                             da, sc, nrun, part);
        else
          k_pc_grid<T, SP><<<grid_pg, kBlock, 0, st>>>(pg, z, odd ? a : a2, acur, w, dn, da, sc,
                                                       nrun, part);
        pp = partials<SP * 2>(st, grid_pg, nr);
      }
    }
    if (!use_pg) {
      k_pc_node<T, SP, true><<<grid_n, kBlock, 0, st>>>(G, z, a, dn, help, sc, nrun);
      k_pc_edge_iter<T, SP><<<grid_e, kBlock, 0, st>>>(G, z, a, w, da, help, sc, nrun, part);
      pp = partials<SP * 2>(st, grid_e, nr);
    }
    s_alpha<SP><<<1, kScalarThreads, 0, st>>>(pp, nr, sc, nrun);
    T2* s_next = s;  // s_k
    if constexpr (kLazy) {
      s_next = sr[(c + 1) & (ring - 1)];
      bool formed = false;  // w formed from help in the pass (after k_pc_grid)
      if constexpr (pg_width(SP)) {
        if (use_pg) {
          k_xr_fused<T, SP, true, true><<<grid_u, kBlock, 0, st>>>(
              nullptr, nullptr, nullptr, r, acur, sr[c & (ring - 1)], s_next, help, nullptr, sm,
              sc, nrun,
              part, nb, wrec());
          formed = true;
        }
      }
      if (!formed)
        k_xr_fused<T, SP, true><<<grid_u, kBlock, 0, st>>>(nullptr, nullptr, nullptr, r, acur,
                                                           sr[c & (ring - 1)], s_next, w, nullptr,
                                                           sm, sc,
                                                           nrun, part, nb);
    } else {
      k_xr_fused<T, SP><<<grid_u, kBlock, 0, st>>>(odd ? x : x2, odd ? x2 : x, p, r, acur, s, s,
                                                   w, ym_or_null(), sm, sc, nrun, part, nb);
    }
    pp = partials<SP * 3>(st, grid_u, nr);
    s_mrs_stop<SP><<<1, kScalarThreads, 0, st>>>(pp, nr, sc, nrun);
    if constexpr (kLazy) {
      // every eighth iteration (k = 8, 16, ...: positions 7 and 15), once the stops up to
      // k - 1 are known: x and p of iterations k-7..k from s_{k-8}..s_{k-1}
      if ((c & (xp - 1)) == xp - 1) {
        const int c0 = (c - (xp - 1)) & (ring - 1);
        XpRing rs{};
        for (int i = 0; i < xp; ++i) rs.s[i] = sr[(c0 + i) & (ring - 1)];
        if (xp == 16)
          k_flush_xp<SP, 16><<<grid_u, kBlock, 0, st>>>(x, p, rs, sc, nrun, nb);
        else
          k_flush_xp<SP, 8><<<grid_u, kBlock, 0, st>>>(x, p, rs, sc, nrun, nb);
      }
    }
    spmv<0>(st, s_next, z, (const T2*)nullptr, (T2*)nullptr, (T2*)nullptr);
    pp = partials<SP * 2>(st, grid_s, nr);
    s_beta<SP, false><<<1, kScalarThreads, 0, st>>>(pp, nr, sc, nrun, maxiter);
  }
};

int ensure(mtcocr_handle* h, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return 0;
  drop_graph(h);  // captured kernels hold the old pointers
  HIPC(b.alloc(bytes));
  return 0;
}

int alloc_work(mtcocr_handle* h, int prec, int sp) {
  const size_t es = (prec == MTCOCR_PREC_FP64) ? sizeof(double2) : sizeof(float2);
  const size_t nv = (size_t)h->nb * sp * es;
  const size_t nm = (size_t)std::max(h->m, 1) * sp * es;
  int rc = 0;
  DevBuf* vecs[] = {&h->vr, &h->vp, &h->vs, &h->vz, &h->va, &h->vw, &h->vym, &h->vsm, &h->vx,
                    &h->vx2, &h->vb};
  for (DevBuf* b : vecs)
    if ((rc = ensure(h, *b, nv))) return rc;
  // the second a buffer of the fused grid preconditioner (batches of 8 or 16 systems; wider
  // FP64 batches run in groups of 16)
  if (sp >= 8 && h->pg_ok && (rc = ensure(h, h->va2, (size_t)h->nb * std::min(sp, 16) * es)))
    return rc;
  // the FP64 solver's ring of s buffers (Engine::sr): 16 for an x/p period of 8; 32 for 16
  // when asked for, or by default when the 16 more vectors leave the device 8 GiB and the
  // compaction buffers (five vectors' worth) free
  if (prec == MTCOCR_PREC_FP64) {
    for (int i = 0; i < 15; ++i)
      if ((rc = ensure(h, h->vsr[i], nv))) return rc;
    int period = h->xp_request;
    if (period == 0) {
      period = 8;
      if (h->vsr[15].bytes >= nv) {
        period = 16;  // held from an earlier solve
      } else {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr > 21 * nv + (size_t(8) << 30))
          period = 16;
      }
    }
    if (period == 16)
      for (int i = 15; i < kRingMax - 1; ++i)
        if ((rc = ensure(h, h->vsr[i], nv))) return rc;
    if (period != h->xp_period) drop_graph(h);  // captured chunks have the old length
    h->xp_period = period;
  }
  if ((rc = ensure(h, h->vhelp, nm))) return rc;
  if ((rc = ensure(h, h->vdn, nm))) return rc;
  h->work_prec = prec;
  h->work_sp = sp;
  return 0;
}

bool use_graphs() {
  const char* e = std::getenv("MTCOCR_NO_GRAPH");
  return !(e && e[0] == '1');
}

// Run COCR iterations until every system of the engine has left the RUNNING state.
template <class T, int SP>
int run_iterations(mtcocr_handle* h, Engine<T, SP>& E, int maxiter, int stop_below = 0) {
  hipStream_t st = h->stream;
  const int prec = std::is_same<T, double>::value ? 0 : 1;
  // 2P iterations per graph in FP64 (P = the x/p period, 8 or 16): the solver's ring of s
  // buffers and its x/p flushes (at positions P - 1 and 2P - 1) are laid out for chunks of 2P
  // that start at iterations 1, 2P + 1, ... (so compaction, between chunks, finds x and p
  // complete and s_k in sr[0]); 8 in FP32
  const int chunk = Engine<T, SP>::kLazy ? 2 * h->xp_period : 8;
  if (use_graphs()) {
    if (!(h->graph && h->graph_prec == prec && h->graph_sp == SP && h->graph_chunk == chunk &&
          h->graph_maxiter == maxiter && h->graph_pg == int(E.use_pg))) {
      if (h->graph) {
        (void)hipGraphExecDestroy(h->graph);
        h->graph = nullptr;
      }
      hipGraph_t g;
      HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int c = 0; c < chunk; ++c) E.iteration(st, maxiter, c);
      HIPC(hipStreamEndCapture(st, &g));
      hipError_t e = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPC(e);
      h->graph_prec = prec;
      h->graph_sp = SP;
      h->graph_chunk = chunk;
      h->graph_maxiter = maxiter;
      h->graph_pg = int(E.use_pg);
    }
  }
  // every running system leaves RUNNING after at most maxiter + 1 iterations; the bound below
  // only guards against a device that stopped executing the kernels
  const long long cap = (long long)maxiter + 2LL * chunk + 16;
  long long launched = 0;
  for (;;) {
    HIPC(hipMemcpyAsync(h->h_flags, E.nrun, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (h->h_flags[0] == 0) break;
    if (h->h_flags[0] <= stop_below) return 1;  // the caller narrows the batch
    if (launched > cap) {
      std::snprintf(g_err, sizeof(g_err), "solver did not retire its systems after %lld iterations",
                    launched);
      return MTCOCR_ERR_HIP;
    }
    if (use_graphs()) {
      HIPC(hipGraphLaunch(h->graph, st));
    } else {
      for (int c = 0; c < chunk; ++c) E.iteration(st, maxiter, c);
    }
    HIPC(hipGetLastError());
    launched += chunk;
  }
  return 0;
}

void set_eps_status(mtcocr_handle* h, std::vector<SysScalars>& hs, int nsys, int sp, double eps) {
  hs.assign(sp, SysScalars());
  for (int s = 0; s < sp; ++s) {
    std::memset(&hs[s], 0, sizeof(SysScalars));
    hs[s].status = (s < nsys) ? ST_RUNNING : ST_INACTIVE;
    hs[s].outer_status = hs[s].status;
    hs[s].eps = eps;
    hs[s].rel = 1.0;
  }
  (void)h;
}

// Outcome of each system of an FP64 batch (caller's batch order), recorded when it leaves.
struct Fp64Run {
  int sp0 = 0;
  std::vector<SysScalars> final_sc;
  std::vector<char> done;
};

template <class F>
int with_sp(int sp, F&& f);

// Widest batch the FP64 solver iterates at once; wider batches run as consecutive groups.
constexpr int kGroupWidth = 16;

// Runs the FP64 iterations of a batch of SP slots (slot j holds the caller's system orig[j]).
// When the running systems fit into half the width, the finished ones are written to x64,
// the running ones are repacked into a batch of the next power of two and the iterations
// continue there: systems that converge late no longer pay for the width of the whole batch.
// Iteration counters, MRS state and x parity move with the systems; D_N^-1 is recomputed for
// the new order. Results are those of the wide batch up to the regrouped dot products.
template <int SP>
int fp64_phase(mtcocr_handle* h, int maxiter, const std::vector<int>& orig,
               const std::vector<double>& om, Fp64Run& R) {
  hipStream_t st = h->stream;
  int rc;
  Engine<double, SP> E(h);
  const int narrowed = run_iterations<double, SP>(h, E, maxiter, SP > 1 ? SP / 2 : 0);
  if (narrowed < 0) return narrowed;
  std::vector<SysScalars> hs(SP);
  HIPC(hipMemcpyAsync(hs.data(), E.sc, sizeof(SysScalars) * SP, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  std::vector<int> sel(SP, -1), running;
  for (int j = 0; j < SP; ++j) {
    if (orig[j] < 0 || hs[j].status == ST_INACTIVE) continue;
    if (hs[j].status == ST_RUNNING) {
      running.push_back(j);
      continue;
    }
    sel[j] = orig[j];
    R.final_sc[orig[j]] = hs[j];
    R.done[orig[j]] = 1;
  }
  if ((rc = ensure(h, h->cmap, sizeof(int) * 128))) return rc;
  const long long nv = (long long)h->nb * SP;
  const int gv = (int)std::min<long long>(4096, (nv + 255) / 256);
  HIPC(hipMemcpyAsync(h->cmap.p, sel.data(), sizeof(int) * SP, hipMemcpyHostToDevice, st));
  // x holds each finished system's answer (k_flush_xp has applied its steps up to `last`)
  k_scatter_x<SP><<<gv, 256, 0, st>>>(h->x64.as<double2>(), R.sp0, E.x, E.x, E.sc,
                                      h->cmap.as<int>(), h->nb);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(st));
  if (narrowed == 0 || running.empty()) return 0;
  // narrow the batch
  const int sp2 = pad_sp((int)running.size());
  std::vector<int> map(sp2, -1), orig2(sp2, -1);
  std::vector<double> om2(sp2, 0.0);
  std::vector<SysScalars> sc2(sp2);
  for (int k = 0; k < sp2; ++k) {
    std::memset(&sc2[k], 0, sizeof(SysScalars));
    sc2[k].status = sc2[k].outer_status = ST_INACTIVE;
  }
  for (int k = 0; k < (int)running.size(); ++k) {
    map[k] = running[k];
    orig2[k] = orig[running[k]];
    om2[k] = om[running[k]];
    sc2[k] = hs[running[k]];
  }
  DevBuf* mains[] = {&h->vx, &h->vx2, &h->vr, &h->vp, &h->vs, &h->vz, &h->va, &h->vw,
                     &h->vym, &h->vsm};
  DevBuf* spares[] = {&h->cvx, &h->cvx2, &h->cvr, &h->cvp, &h->cvs, &h->cvz, &h->cva, &h->cvw,
                      &h->cvym, &h->cvsm};
  HIPC(hipMemcpyAsync(h->cmap.p, map.data(), sizeof(int) * sp2, hipMemcpyHostToDevice, st));
  const long long nv2 = (long long)h->nb * sp2;
  const int g2 = (int)std::min<long long>(4096, (nv2 + 255) / 256);
  for (int i = 0; i < 10; ++i) {
    if ((rc = ensure(h, *spares[i], sizeof(double2) * (size_t)nv2))) return rc;
    k_repack<double2><<<g2, 256, 0, st>>>(mains[i]->as<double2>(), SP, spares[i]->as<double2>(),
                                          sp2, h->cmap.as<int>(), h->nb);
  }
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(st));
  for (int i = 0; i < 10; ++i) {
    std::swap(mains[i]->p, spares[i]->p);
    std::swap(mains[i]->bytes, spares[i]->bytes);
  }
  drop_graph(h);  // captured iterations hold the old buffers and width
  const int nrun = (int)running.size();
  HIPC(hipMemcpyAsync(h->scal.p, sc2.data(), sizeof(SysScalars) * sp2, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(h->omega.p, om2.data(), sizeof(double) * sp2, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(h->nrun.p, &nrun, sizeof(int), hipMemcpyHostToDevice, st));
  HIPC(hipStreamSynchronize(st));
  return with_sp(sp2, [&](auto spc) -> int {
    constexpr int SP2 = decltype(spc)::value;
    if constexpr (SP2 < SP) {
      Engine<double, SP2> E2(h);
      E2.setup_diagonals(st);
      HIPC(hipGetLastError());
      return fp64_phase<SP2>(h, maxiter, orig2, om2, R);
    }
    return MTCOCR_ERR_STATE;
  });
}

// Plain FP64 COCR on the systems of `mask` (all when null), warm-started from x64.
template <int SP>
int solve_fp64(mtcocr_handle* h, double eps, int maxiter, mtcocr_stats* stats,
               const std::vector<char>* mask = nullptr) {
  hipStream_t st = h->stream;
  int rc = alloc_work(h, MTCOCR_PREC_FP64, SP);
  if (rc) return rc;
  Engine<double, SP> E(h);
  std::vector<SysScalars> hs;
  set_eps_status(h, hs, h->nsys, SP, eps);
  if (mask)
    for (int s = 0; s < h->nsys; ++s)
      if (!(*mask)[s]) hs[s].status = hs[s].outer_status = ST_INACTIVE;
  std::vector<double> om0(SP);
  HIPC(hipMemcpyAsync(om0.data(), h->omega.p, sizeof(double) * SP, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(E.sc, hs.data(), sizeof(SysScalars) * SP, hipMemcpyHostToDevice, st));
  int one = 1;
  HIPC(hipMemcpyAsync(E.nrun, &one, sizeof(int), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(E.x, h->x64.p, (size_t)h->nb * SP * sizeof(double2), hipMemcpyDeviceToDevice, st));
  HIPC(hipMemcpyAsync(E.b, h->b64.p, (size_t)h->nb * SP * sizeof(double2), hipMemcpyDeviceToDevice, st));
  HIPC(hipStreamSynchronize(st));
  // keep the wider set of work vectors as the main one (compaction swaps in narrower ones;
  // their contents are spent once the answers are in x64)
  auto keep_wider = [&]() {
    DevBuf* mains[] = {&h->vx, &h->vx2, &h->vr, &h->vp, &h->vs, &h->vz, &h->va, &h->vw,
                       &h->vym, &h->vsm};
    DevBuf* spares[] = {&h->cvx, &h->cvx2, &h->cvr, &h->cvp, &h->cvs, &h->cvz, &h->cva,
                        &h->cvw, &h->cvym, &h->cvsm};
    for (int i = 0; i < 10; ++i)
      if (spares[i]->bytes > mains[i]->bytes) {
        std::swap(mains[i]->p, spares[i]->p);
        std::swap(mains[i]->bytes, spares[i]->bytes);
        drop_graph(h);
      }
  };
  Fp64Run R;
  R.sp0 = SP;
  R.final_sc.assign(SP, SysScalars());
  R.done.assign(SP, 0);
  // timed span: DFP setup (diagonals) + COCR loop, like the reference's kit interval
  HIPC(hipEventRecord(h->ev0, st));
  if constexpr (SP <= kGroupWidth) {
    E.setup_diagonals(st);
    E.prologue(st, maxiter);
    HIPC(hipGetLastError());
    std::vector<int> orig(SP, -1);
    for (int s = 0; s < h->nsys; ++s) orig[s] = s;
    rc = fp64_phase<SP>(h, maxiter, orig, om0, R);
    if (rc) return rc;
    keep_wider();
  } else {
    // A wide batch runs as consecutive groups of kGroupWidth neighbouring systems: each group's
    // x and b are repacked out of the caller's [row][system] arrays, the group runs (and
    // narrows) on its own, and its answers are scattered back to x64. On config-3 models
    // (32 systems) this was 8% faster than one batch of 32.
    constexpr int GW = kGroupWidth;
    if ((rc = ensure(h, h->cmap, sizeof(int) * 128))) return rc;
    for (int g0 = 0; g0 < h->nsys; g0 += GW) {
      const int cnt = std::min(GW, h->nsys - g0);
      std::vector<int> map(GW, -1), orig(GW, -1);
      std::vector<double> om(GW, 0.0);
      std::vector<SysScalars> sc(GW);
      for (int k = 0; k < GW; ++k) {
        std::memset(&sc[k], 0, sizeof(SysScalars));
        sc[k].status = sc[k].outer_status = ST_INACTIVE;
      }
      int active = 0;
      for (int k = 0; k < cnt; ++k) {
        map[k] = g0 + k;
        om[k] = om0[g0 + k];
        sc[k] = hs[g0 + k];
        if (hs[g0 + k].status == ST_RUNNING) {
          orig[k] = g0 + k;
          ++active;
        }
      }
      if (active == 0) continue;
      Engine<double, GW> G(h);  // after keep_wider: the main vectors hold GW systems or more
      const long long nv = (long long)h->nb * GW;
      const int gv = (int)std::min<long long>(4096, (nv + 255) / 256);
      HIPC(hipMemcpyAsync(h->cmap.p, map.data(), sizeof(int) * GW, hipMemcpyHostToDevice, st));
      k_repack<double2><<<gv, 256, 0, st>>>(h->x64.as<double2>(), SP, G.x, GW,
                                            h->cmap.as<int>(), h->nb);
      k_repack<double2><<<gv, 256, 0, st>>>(h->b64.as<double2>(), SP, G.b, GW,
                                            h->cmap.as<int>(), h->nb);
      HIPC(hipMemcpyAsync(h->scal.p, sc.data(), sizeof(SysScalars) * GW, hipMemcpyHostToDevice,
                          st));
      HIPC(hipMemcpyAsync(h->omega.p, om.data(), sizeof(double) * GW, hipMemcpyHostToDevice, st));
      HIPC(hipMemcpyAsync(h->nrun.p, &one, sizeof(int), hipMemcpyHostToDevice, st));
      drop_graph(h);  // a captured graph holds the previous width
      G.setup_diagonals(st);
      G.prologue(st, maxiter);
      HIPC(hipGetLastError());
      HIPC(hipStreamSynchronize(st));  // the host vectors above stay alive until here
      rc = fp64_phase<GW>(h, maxiter, orig, om, R);
      if (rc) return rc;
      keep_wider();
    }
  }
  HIPC(hipEventRecord(h->ev1, st));
  HIPC(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  // the batch in the caller's order again, true residual of x64 in FP64
  HIPC(hipMemcpyAsync(h->omega.p, om0.data(), sizeof(double) * SP, hipMemcpyHostToDevice, st));
  std::vector<SysScalars> tmp(hs);
  for (int s = 0; s < SP; ++s) tmp[s].status = (s < h->nsys) ? ST_RUNNING : ST_INACTIVE;
  HIPC(hipMemcpyAsync(E.sc, tmp.data(), sizeof(SysScalars) * SP, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(E.nrun, &one, sizeof(int), hipMemcpyHostToDevice, st));
  E.template spmv<1>(st, h->x64.as<double2>(), h->r64.as<double2>(), h->b64.as<double2>(),
                     (double2*)nullptr, (double2*)nullptr);
  HIPC(hipGetLastError());
  std::vector<double> partials((size_t)E.grid_s * SP * 2);
  HIPC(hipMemcpyAsync(partials.data(), E.part, partials.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  for (int s = 0; s < h->nsys; ++s) {
    if (mask && !(*mask)[s]) continue;
    double bb = 0, rr = 0;
    for (int b = 0; b < E.grid_s; ++b) {
      bb += partials[((size_t)b * SP + s) * 2];
      rr += partials[((size_t)b * SP + s) * 2 + 1];
    }
    const SysScalars& f = R.final_sc[s];
    mtcocr_stats& o = stats[s];
    o.req = !R.done[s] ? MTCOCR_ERR_STATE : f.status;
    o.iter = f.iter;
    o.outer = 1;
    o.flags = 0;
    o.residual_rel = f.rel;
    o.true_rel_res = bb > 0 ? std::sqrt(rr) / std::sqrt(bb) : 0.0;
    o.seconds = ms * 1e-3;
  }
  return 0;
}

template <int SP>
int solve_mixed(mtcocr_handle* h, double eps, int maxiter, mtcocr_stats* stats) {
  hipStream_t st = h->stream;
  int rc = alloc_work(h, MTCOCR_PREC_MIXED, SP);
  if (rc) return rc;
  std::vector<SysScalars> hs;
  float ms = 0.f;
  {
    Engine<float, SP> E(h);
    set_eps_status(h, hs, h->nsys, SP, eps);
    for (int s = 0; s < SP; ++s) hs[s].stall_window = 200;
    HIPC(hipMemcpyAsync(E.sc, hs.data(), sizeof(SysScalars) * SP, hipMemcpyHostToDevice, st));
    int one = 1;
    int* nrun_outer = h->nrun2.as<int>();
    HIPC(hipMemcpyAsync(nrun_outer, &one, sizeof(int), hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
    // Each FP32 pass aims at most two orders of magnitude below the residual it starts from.
    // In the air the FP32 vectors stall well above eps, so the FP64 finish does the rest: on
    // the config-2 batch a 1e-2 floor took 0.22 s per system, 1e-3 0.25 s and 1e-5 0.30 s.
    const double inner_floor = 1e-2;
    const int max_outer = 40;
    HIPC(hipEventRecord(h->ev0, st));
    E.setup_diagonals(st);
    const long long nv = (long long)h->nb * SP;
    const int gv = (int)std::min<long long>(4096, (nv + 255) / 256);
    for (int pass = 0;; ++pass) {
      if (pass > max_outer + 1) {
        std::snprintf(g_err, sizeof(g_err), "refinement did not finish after %d passes", pass);
        return MTCOCR_ERR_HIP;
      }
      // FP64 true residual r64 = b - A x64 on the systems still refining
      launch_spmv<double, SP, 2>(h, E.layout, E.grid_s, st, E.omega, h->x64.as<double2>(),
                                 h->r64.as<double2>(), h->b64.as<double2>(), nullptr, nullptr,
                                 E.sc, nrun_outer, E.part);
      int nr;
      const double* pp = E.template partials<SP * 2>(st, E.grid_s, nr);
      s_outer<SP><<<1, kScalarThreads, 0, st>>>(pp, nr, E.sc, nrun_outer, E.nrun, eps,
                                                inner_floor, pass == 0 ? 1 : 0, max_outer);
      HIPC(hipMemcpyAsync(h->h_flags, nrun_outer, sizeof(int), hipMemcpyDeviceToHost, st));
      HIPC(hipStreamSynchronize(st));
      k_rollback<SP><<<gv, 256, 0, st>>>(h->x64.as<double2>(), E.ym, E.sc, h->nb);
      if (h->h_flags[0] == 0) break;
      k_to_inner<SP><<<gv, 256, 0, st>>>(h->r64.as<double2>(), E.b, E.x, E.sc, h->nb);
      E.prologue(st, maxiter);
      HIPC(hipGetLastError());
      rc = run_iterations<float, SP>(h, E, maxiter);
      if (rc) return rc;
      s_inner_done<SP><<<1, 64, 0, st>>>(E.sc);
      k_from_inner<SP><<<gv, 256, 0, st>>>(h->x64.as<double2>(), E.ym, E.sc, h->nb);
    }
    HIPC(hipEventRecord(h->ev1, st));
    HIPC(hipEventSynchronize(h->ev1));
    HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    HIPC(hipMemcpyAsync(hs.data(), E.sc, sizeof(SysScalars) * SP, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
  }
  std::vector<char> fall(h->nsys, 0);
  bool any = false;
  for (int s = 0; s < h->nsys; ++s) {
    mtcocr_stats& o = stats[s];
    o.req = hs[s].outer_status == ST_RUNNING ? MTCOCR_ERR_NOTCONV : hs[s].outer_status;
    o.iter = hs[s].total_iter;
    o.outer = hs[s].outer;
    o.flags = 0;
    o.residual_rel = hs[s].rel;
    o.true_rel_res = hs[s].rel;
    o.seconds = ms * 1e-3;
    if (o.req == MTCOCR_ERR_NOTCONV) {
      fall[s] = 1;
      any = true;
    }
  }
  // FP64 finish of the systems the refinement did not converge, warm-started from x64. Its MRS
  // stop test runs on the smoothed residual, which can sit several times below the true one:
  // a system whose true residual still misses eps gets a second, tighter finish.
  double extra = 0.0;
  for (int attempt = 0; attempt < 2 && any; ++attempt) {
    std::vector<mtcocr_stats> fs(h->nsys);
    rc = solve_fp64<SP>(h, attempt == 0 ? eps / 3.0 : eps / 30.0, maxiter, fs.data(), &fall);
    if (rc) return rc;
    any = false;
    bool counted = false;
    for (int s = 0; s < h->nsys; ++s) {
      if (!fall[s]) continue;
      mtcocr_stats& o = stats[s];
      o.iter += fs[s].iter;
      o.flags |= MTCOCR_FLAG_FP64_FINISH;
      o.residual_rel = fs[s].residual_rel;
      o.true_rel_res = fs[s].true_rel_res;
      if (!counted) {
        extra += fs[s].seconds;
        counted = true;
      }
      if (fs[s].req == MTCOCR_OK && fs[s].true_rel_res >= eps && attempt == 0) {
        o.req = MTCOCR_ERR_NOTCONV;  // unless the next attempt gets there
        any = true;
      } else {
        // after the second attempt the FP64 stop test (at eps/30) decides, as for the plain
        // FP64 solver; a true residual still at or above eps is flagged
        o.req = fs[s].req;
        if (fs[s].req == MTCOCR_OK && fs[s].true_rel_res >= eps) o.flags |= MTCOCR_FLAG_RES_FLOOR;
        fall[s] = 0;
      }
    }
  }
  for (int s = 0; s < h->nsys; ++s) stats[s].seconds += extra;
  return 0;
}

template <class F>
int with_sp(int sp, F&& f) {
  switch (sp) {
    case 1: return f(std::integral_constant<int, 1>());
    case 2: return f(std::integral_constant<int, 2>());
    case 4: return f(std::integral_constant<int, 4>());
    case 8: return f(std::integral_constant<int, 8>());
    case 16: return f(std::integral_constant<int, 16>());
    case 32: return f(std::integral_constant<int, 32>());
    case 64: return f(std::integral_constant<int, 64>());
  }
  return MTCOCR_ERR_ARG;
}

int set_batch(mtcocr_handle* h, int nsys, const double* omega) {
  if (nsys < 1 || nsys > 64) return MTCOCR_ERR_ARG;
  const int sp = pad_sp(nsys);
  std::vector<double> om(sp, 0.0);
  for (int s = 0; s < nsys; ++s) om[s] = omega ? omega[s] : 1.0;
  int rc;
  if ((rc = ensure(h, h->omega, sizeof(double) * 64))) return rc;
  if ((rc = ensure(h, h->scal, sizeof(SysScalars) * 64))) return rc;
  if ((rc = ensure(h, h->nrun, 64))) return rc;
  if ((rc = ensure(h, h->nrun2, 64))) return rc;
  if ((rc = ensure(h, h->part, sizeof(double) * (size_t)kMaxBlocks * 64 * 3))) return rc;
  if ((rc = ensure(h, h->part2, sizeof(double) * (size_t)(kMaxBlocks / kFold) * 64 * 3)))
    return rc;
  const size_t v64 = (size_t)h->nb * sp * sizeof(double2);
  if ((rc = ensure(h, h->b64, v64))) return rc;
  if ((rc = ensure(h, h->x64, v64))) return rc;
  if ((rc = ensure(h, h->r64, v64))) return rc;
  HIPC(hipMemcpyAsync(h->omega.p, om.data(), sizeof(double) * sp, hipMemcpyHostToDevice, h->stream));
  h->nsys = nsys;
  h->sp = sp;
  return 0;
}

// interleave nsys host vectors (each 2*nb doubles) into the [row][system] device layout
// (device row r holds the caller's row order[r])
int put_batch(mtcocr_handle* h, DevBuf& dst, const double* src) {
  const int sp = h->sp, nb = h->nb;
  const int* ord = h->order.empty() ? nullptr : h->order.data();
  std::vector<double2> tmp((size_t)nb * sp);
  for (int r = 0; r < nb; ++r)
    for (int s = 0; s < sp; ++s) {
      double2 v;
      const size_t o = ord ? (size_t)ord[r] : (size_t)r;
      if (s < h->nsys && src) {
        v.x = src[(size_t)s * 2 * nb + 2 * o];
        v.y = src[(size_t)s * 2 * nb + 2 * o + 1];
      } else {
        v.x = 0.0;
        v.y = 0.0;
      }
      tmp[(size_t)r * sp + s] = v;
    }
  HIPC(hipMemcpyAsync(dst.p, tmp.data(), tmp.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return 0;
}

int get_batch(mtcocr_handle* h, const DevBuf& srcb, double* dst) {
  const int sp = h->sp, nb = h->nb;
  std::vector<double2> tmp((size_t)nb * sp);
  HIPC(hipMemcpyAsync(tmp.data(), srcb.p, tmp.size() * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  const int* ord = h->order.empty() ? nullptr : h->order.data();
  for (int r = 0; r < nb; ++r) {
    const size_t o = ord ? (size_t)ord[r] : (size_t)r;
    for (int s = 0; s < h->nsys; ++s) {
      dst[(size_t)s * 2 * nb + 2 * o] = tmp[(size_t)r * sp + s].x;
      dst[(size_t)s * 2 * nb + 2 * o + 1] = tmp[(size_t)r * sp + s].y;
    }
  }
  return 0;
}

int upload_matrix(mtcocr_handle* h, const SlicedEll& E, bool ks) {
  hipStream_t st = h->stream;
  int rc;
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  h->have_matrix = false;
  h->nb = E.nb;
  h->nslices = E.nslices;
  h->nnz_c = E.nnz_complex;
  h->nnz_r = E.nnz_real;
  h->ks_form = ks;
  if ((rc = upload(h, h->offc, E.offc, st)) || (rc = upload(h, h->nc, E.nc, st)) ||
      (rc = upload(h, h->offr, E.offr, st)) || (rc = upload(h, h->nr, E.nr, st)) ||
      (rc = upload(h, h->colc, E.colc, st)) || (rc = upload(h, h->colr, E.colr, st)) ||
      (rc = upload(h, h->ac64, E.ac, st)) || (rc = upload(h, h->bc64, E.bc, st)) ||
      (rc = upload(h, h->ar64, E.ar, st)) || (rc = upload(h, h->da64, E.diag_a, st)) ||
      (rc = upload(h, h->db64, E.diag_b, st)) || (rc = upload(h, h->rp, E.rp, st)) ||
      (rc = upload(h, h->rb, E.rb, st)) || (rc = upload(h, h->pcol, E.pcol, st)) ||
      (rc = upload(h, h->pa, E.pa, st)) || (rc = upload(h, h->pb, E.pb, st)) ||
      (rc = upload(h, h->hoff, E.hoff, st)) || (rc = upload(h, h->halo, E.halo, st)) ||
      (rc = upload(h, h->lcol, E.lcol, st)) || (rc = upload(h, h->ldiag, E.ldiag, st)))
    return rc;
  h->halo_max = E.hoff.empty() ? 0 : E.halo_max;
  h->nblk = E.hoff.empty() ? 0 : int(E.hoff.size()) - 1;
  h->stc_ok = false;
  h->stc.release();
  h->stc_big = false;
  if (h->grid_order) {
    std::vector<double> coef;
    long long box = 0;  // the largest direction box's coefficient bytes
    for (int d = 0; d < 3; ++d) box = std::max<long long>(box, h->grid.count(d) * kStcCoef * 8);
    if (stencil_from_rows(E, h->grid, coef) == 0 && box <= (long long)kOob) {
      if ((rc = upload(h, h->stc, coef, st))) return rc;
      h->stc_ok = true;
      h->stc_big = (long long)E.nb * kStcCoef * 8 > (long long)kOob;
    }
  }
  HIPC(hipStreamSynchronize(st));
  h->have_matrix = true;
  // a gradient built for another edge count does not fit this matrix: set it again
  if (h->have_g && h->g_nb != h->nb) h->have_g = false;
  h->nsys = 0;
  h->work_prec = -1;
  return 0;
}

// The tuner for one batch width sp (the first min(nsys, sp) systems of the batch on the handle);
// it stores the chosen shapes in the handle's per-width tables and returns the SpMV time at the
// chosen shape in *spmv_ms_out.
int autotune_width(mtcocr_handle* h, int precision, int reps, int sp, double* spmv_ms_out) {
  int rc = alloc_work(h, precision, sp);
  if (rc) return rc;
  hipStream_t st = h->stream;
  std::vector<SysScalars> hs;
  set_eps_status(h, hs, std::min(h->nsys, sp), sp, 1e-8);
  for (auto& q : hs) q.mode = MRS_BLEND;  // the update pass's common step
  HIPC(hipMemcpyAsync(h->scal.p, hs.data(), sizeof(SysScalars) * sp, hipMemcpyHostToDevice, st));
  int one = 1;
  HIPC(hipMemcpyAsync(h->nrun.p, &one, sizeof(int), hipMemcpyHostToDevice, st));
  // large grids: the streaming kernels reach their best bandwidth with about one 256-element
  // tile per workgroup (the resident workgroups then sweep neighbouring tiles together)
  static const int kCands[] = {512, 1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072};
  const int ki = mtcocr_handle::sp_index(sp), q = precision;
  int* caps[4] = {&h->cap_spmv[q][ki], &h->cap_node[q][ki], &h->cap_vec[q][ki],
                  &h->cap_upd[q][ki]};
  int best[4] = {*caps[0], *caps[1], *caps[2], *caps[3]};
  int best_layout = h->spmv_layout[q][ki];
  double best_ms[4] = {1e30, 1e30, 1e30, 1e30};
  rc = with_sp(sp, [&](auto spc) -> int {
    constexpr int SP = decltype(spc)::value;
    auto run = [&](auto tag) -> int {
      using T = decltype(tag);
      using T2 = typename Vec2<T>::type;
      {
        Engine<T, SP> E(h);
        const long long nv = (long long)E.nb * SP, nm = (long long)std::max(E.m, 1) * SP;
        const int gf = (int)std::min<long long>(2048, (nv + 255) / 256);
        const int gm = (int)std::min<long long>(2048, (nm + 255) / 256);
        for (T2* v : {E.s, E.z, E.a, E.p, E.w, E.x, E.x2, E.r, E.ym, E.sm, E.sr[1]})
          k_fill<T2><<<gf, 256, 0, st>>>(v, nv);
        k_fill<T2><<<gm, 256, 0, st>>>(E.help, nm);
        E.setup_diagonals(st);
        HIPC(hipGetLastError());
      }
      for (int fam = 0; fam < 4; ++fam) {
        // SpMV layouts that apply to this batch (the others fall back to one of these)
        std::vector<int> layouts{0};
        if (fam == 0 && SP >= 4 && rows_fit(h, SP, int(sizeof(T2)))) {
          layouts.push_back(1);
          if (SP <= 16 && blk_lds_bytes<T>(h) > 0) layouts.push_back(2);
          if (SP == 16 && h->stc_ok) {
            if (!h->stc_big) layouts.push_back(3);
            layouts.push_back(4);
          }
          if (SP == 8 && h->stc_ok) layouts.push_back(4);
        }
        int best_kc = h->stc_kc[q][ki];
        int best_band = h->stl_band[q][ki];
        for (int layout : layouts) {
          // the stencil kernel's sweep length: a few lengths, each at every cap
          std::vector<int> kcs{0};
          if (layout >= 3) {
            kcs = {2, 4, 8, 16, h->grid.nz};
            if (layout == 4) kcs = {8, 16, (h->grid.nz + 1) / 2, h->grid.nz};
            std::sort(kcs.begin(), kcs.end());
            kcs.erase(std::unique(kcs.begin(), kcs.end()), kcs.end());
          }
          // the LDS-staged kernel's task order: bands of tile columns (0: one band)
          std::vector<int> bands{h->stl_band[q][ki]};
          if (layout == 4) bands = {0, 5, 9, 13};
          for (int kc : kcs) {
          for (int band : bands) {
          if (layout >= 3) h->stc_kc[q][ki] = kc;
          if (layout == 4) h->stl_band[q][ki] = band;
          // candidates: fixed caps, plus the kernel's resident capacity (a grid that runs in
          // one wave, whose tile walk then sweeps each XCD's rows as one contiguous front) and
          // twice that
          int per_cu = 0;
          if (fam == 0 && layout == 3) {
            if constexpr (SP == 16)
              HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, reinterpret_cast<const void*>(k_spmv_stc<T, SP, 0>), kBlock, 0));
          } else if (fam == 0 && layout == 4) {
            if constexpr (SP == 16 || SP == 8)
              HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, reinterpret_cast<const void*>(k_spmv_stl<T, SP, 0>), kBlock, 0));
          } else if (fam == 0 && layout == 2) {
            if constexpr (SP >= 4 && SP <= 16)
              HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, reinterpret_cast<const void*>(k_spmv_blk<T, SP, 0>), kBlock,
                  blk_lds_bytes<T>(h)));
          } else if (fam == 0 && layout == 1) {
            if constexpr (SP >= 4)
              HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, reinterpret_cast<const void*>(k_spmv_rg<T, SP, 0>), kBlock, 0));
          } else if (fam == 0) {
            HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(k_spmv<T, SP, 0>), kBlock, 0));
          } else if (fam == 1) {
            HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(k_pc_node<T, SP, true>), kBlock, 0));
          } else if (fam == 2) {
            HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(k_pc_edge_iter<T, SP>), kBlock, 0));
          } else {
            HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu,
                reinterpret_cast<const void*>(k_xr_fused<T, SP, Engine<T, SP>::kLazy>),
                kBlock, 0));
          }
          std::vector<int> cands(std::begin(kCands), std::end(kCands));
          const int resident = round_up(std::max(1, per_cu * h->n_cu), kXcd);
          for (int c : {resident, 2 * resident})
            if (c <= kMaxBlocks) cands.push_back(c);
          std::sort(cands.begin(), cands.end());
          cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
          int last_grid = -1;
          for (int cap : cands) {
            *caps[fam] = cap;
            if (fam == 0) h->spmv_layout[q][ki] = layout;
            Engine<T, SP> E(h);
            const int grid = fam == 0 ? E.grid_s : fam == 1 ? E.grid_n : fam == 2 ? E.grid_e
                                                                                   : E.grid_u;
            if (grid == last_grid) continue;  // the cap no longer limits this grid
            last_grid = grid;
            // each timed unit is the kernel plus the one-block reduction of its partials
            double* probe_out = h->r64.as<double>();  // scratch: the outer residual buffer
            auto launch = [&]() {
              if (fam == 0) {
                E.template spmv<0>(st, E.s, E.z, (const T2*)nullptr, (T2*)nullptr,
                                   (T2*)nullptr);
                int nr;
                const double* pp = E.template partials<SP * 2>(st, E.grid_s, nr);
                s_reduce_probe<SP * 2><<<1, kScalarThreads, 0, st>>>(pp, nr, probe_out);
              } else if (fam == 1) {
                k_pc_node<T, SP, true><<<E.grid_n, kBlock, 0, st>>>(E.G, E.z, E.a, E.dn, E.help,
                                                                    E.sc, E.nrun);
              } else if (fam == 2) {
                // after the node pass (at its tuned shape), as in an iteration: the edge
                // kernel's gathers of help then find what the node pass left in the caches
                k_pc_node<T, SP, true><<<E.grid_n, kBlock, 0, st>>>(E.G, E.z, E.a, E.dn, E.help,
                                                                    E.sc, E.nrun);
                k_pc_edge_iter<T, SP><<<E.grid_e, kBlock, 0, st>>>(E.G, E.z, E.a, E.w,
                                                                   E.da, E.help, E.sc, E.nrun,
                                                                   E.part);
                int nr;
                const double* pp = E.template partials<SP * 2>(st, E.grid_e, nr);
                s_reduce_probe<SP * 2><<<1, kScalarThreads, 0, st>>>(pp, nr, probe_out);
              } else {
                // the pass without x and p (FP64: s into the ring). After the fused grid
                // preconditioner the solver runs the variant that forms w from help; tuning
                // the cap on that variant alone picked a cap that made the iterations 0.4%
                // slower, so the cap is tuned on this one
                if constexpr (Engine<T, SP>::kLazy)
                  k_xr_fused<T, SP, true><<<E.grid_u, kBlock, 0, st>>>(
                      nullptr, nullptr, nullptr, E.r, E.a, E.s, E.sr[1], E.w, nullptr, E.sm,
                      E.sc, E.nrun, E.part, E.nb);
                else
                  k_xr_fused<T, SP><<<E.grid_u, kBlock, 0, st>>>(
                      E.x, E.x2, E.p, E.r, E.a, E.s, E.s, E.w, E.ym_or_null(), E.sm, E.sc,
                      E.nrun, E.part, E.nb);
                int nr;
                const double* pp = E.template partials<SP * 3>(st, E.grid_u, nr);
                s_reduce_probe<SP * 3><<<1, kScalarThreads, 0, st>>>(pp, nr, probe_out);
              }
            };
            launch();
            HIPC(hipGetLastError());
            HIPC(hipEventRecord(h->ev0, st));
            for (int i = 0; i < reps; ++i) launch();
            HIPC(hipEventRecord(h->ev1, st));
            HIPC(hipGetLastError());
            HIPC(hipEventSynchronize(h->ev1));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
            const double avg = double(ms) / reps;
            if (avg < best_ms[fam]) {
              best_ms[fam] = avg;
              best[fam] = cap;
              if (fam == 0) best_layout = layout;
              if (fam == 0 && layout >= 3) best_kc = kc;
              if (fam == 0 && layout == 4) best_band = band;
            }
          }
          }
          }
        }
        *caps[fam] = best[fam];
        if (fam == 0) {
          h->spmv_layout[q][ki] = best_layout;
          h->stc_kc[q][ki] = best_kc;
          h->stl_band[q][ki] = best_band;
        }
      }
      return 0;
    };
    if (precision == MTCOCR_PREC_FP64) return run(double());
    return run(float());
  });
  drop_graph(h);
  if (rc) return rc;
  *spmv_ms_out = best_ms[0];
  return 0;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C-ABI

extern "C" {

const char* mtcocr_error_string(int code) {
  switch (code) {
    case MTCOCR_OK: return "converged";
    case MTCOCR_X0_NULL: return "initial guess is already the solution";
    case MTCOCR_DIV_BY_ZERO: return "division by zero inside COCR";
    case MTCOCR_MAXITER: return "maximum number of iterations reached";
    case MTCOCR_ERR_ARG: return "invalid argument";
    case MTCOCR_ERR_HIP: return g_err[0] ? g_err : "HIP runtime error";
    case MTCOCR_ERR_ALLOC: return "device memory allocation failed";
    case MTCOCR_ERR_STATE: return "matrix, gradient or right-hand side not set";
    case MTCOCR_ERR_NOTCONV: return "iterative refinement stalled above the tolerance";
  }
  return "unknown error";
}

int mtcocr_create(int device, mtcocr_handle** out) {
  if (!out) return MTCOCR_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  HIPC(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(device));
  std::unique_ptr<mtcocr_handle> h(new mtcocr_handle());
  h->device = device;
  HIPC(hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, device));
  HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPC(hipEventCreate(&h->ev0));
  HIPC(hipEventCreate(&h->ev1));
  HIPC(hipHostMalloc((void**)&h->h_flags, 64, hipHostMallocDefault));
  // the allocation-free entry point's scratch address on this device (a failed lookup only
  // disables that entry point, which then reports it)
  (void)scratch_ptr(device, nullptr);
  *out = h.release();
  return 0;
}

int mtcocr_destroy(mtcocr_handle* h) {
  if (!h) return MTCOCR_ERR_ARG;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->graph) (void)hipGraphExecDestroy(h->graph);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->h_flags) (void)hipHostFree(h->h_flags);
  hipStream_t st = h->stream;
  delete h;  // DevBuf destructors free device memory
  if (st) (void)hipStreamDestroy(st);
  return 0;
}

int mtcocr_set_matrix_block2x2(mtcocr_handle* h, int nb, const int* ig, const int* jg,
                               const int* idi, const int* ijg, const double* di,
                               const double* gg) {
  if (!h || nb <= 0) return MTCOCR_ERR_ARG;
  if (!h->order.empty() && (int)h->order.size() != nb) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(h->device));
  SlicedEll E;
  int rc = ell_from_block2x2(nb, ig, jg, idi, ijg, di, gg,
                             h->order.empty() ? nullptr : h->order.data(), E);
  if (rc) return rc;
  return upload_matrix(h, E, false);
}

int mtcocr_set_matrix_ks(mtcocr_handle* h, int nb, const int* ig, const int* jg,
                         const double* k_di, const double* k_gg, const double* s_di,
                         const double* s_gg) {
  if (!h || nb <= 0) return MTCOCR_ERR_ARG;
  if (!h->order.empty() && (int)h->order.size() != nb) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(h->device));
  SlicedEll E;
  int rc = ell_from_ks(nb, ig, jg, k_di, k_gg, s_di, s_gg,
                       h->order.empty() ? nullptr : h->order.data(), E);
  if (rc) return rc;
  return upload_matrix(h, E, true);
}

int mtcocr_set_gradient_mesh(mtcocr_handle* h, int kpar, int n_edges, int n_nodes_c,
                             const int* nvkat, const int* nver, const double* sigma3d,
                             const int* edges, const int* ig_t, const int* jg_t,
                             const double* gg_t, const int* is_node_bound) {
  if (!h) return MTCOCR_ERR_ARG;
  if (h->have_matrix && n_edges != h->nb) return MTCOCR_ERR_ARG;
  if (!h->order.empty() && (int)h->order.size() != n_edges) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(h->device));
  GradientEll G;
  int rc = build_gradient(kpar, n_edges, n_nodes_c, nvkat, nver, sigma3d, edges, ig_t, jg_t, gg_t,
                          is_node_bound, h->order.empty() ? nullptr : h->order.data(), G);
  if (rc) return rc;
  if (G.wg > kMaxGRow) {
    std::snprintf(g_err, sizeof(g_err),
                  "a node of G has %d entries; the preconditioner set-up takes at most %d", G.wg,
                  kMaxGRow);
    return MTCOCR_ERR_ARG;
  }
  hipStream_t st = h->stream;
  h->have_g = false;
  if ((rc = upload(h, h->gcol, G.gcol, st)) || (rc = upload(h, h->gtcol, G.gtcol, st)) ||
      (rc = upload(h, h->gval64, G.gval, st)) || (rc = upload(h, h->gtval64, G.gtval, st)))
    return rc;
  HIPC(hipStreamSynchronize(st));
  if ((rc = upload_cast_float(h, h->gval32, G.gval, st)) ||
      (rc = upload_cast_float(h, h->gtval32, G.gtval, st)))
    return rc;
  h->m = G.m;
  h->wg = G.wg;
  h->wgt = G.wgt;
  h->g_nb = n_edges;
  h->pg_ok = h->grid_order && gradient_on_grid(G, h->grid);
  drop_graph(h);  // captured iterations hold the preconditioner path
  h->have_g = true;
  h->work_prec = -1;
  return 0;
}

int mtcocr_set_mesh_order(mtcocr_handle* h, int kpar, int n_nodes, const int* nver,
                          int n_edges, const int* edges, int block) {
  if (!h || (block < 0 && block != MTCOCR_ORDER_GRID)) return MTCOCR_ERR_ARG;
  if (h->have_matrix || h->have_g) return MTCOCR_ERR_STATE;  // the layouts are built in order
  h->grid_order = false;
  if (block == 0) {
    h->order.clear();
    return 0;
  }
  std::vector<int> ord;
  GridDims dims;
  const int rc = block == MTCOCR_ORDER_GRID
                     ? grid_order_from_mesh(kpar, n_nodes, nver, n_edges, edges, ord, dims)
                     : order_from_mesh(kpar, n_nodes, nver, n_edges, edges, block, ord);
  if (rc) {
    h->order.clear();
    return rc;
  }
  h->order.swap(ord);
  if (block == MTCOCR_ORDER_GRID) {
    h->grid_order = true;
    h->grid = dims;
  }
  return 0;
}

int mtcocr_set_rhs(mtcocr_handle* h, int nsys, const double* omega, const double* pr,
                   const double* x0) {
  if (!h || !pr) return MTCOCR_ERR_ARG;
  if (!h->have_matrix) return MTCOCR_ERR_STATE;
  if (!h->ks_form && omega) return MTCOCR_ERR_ARG;  // one matrix, several right-hand sides
  if (h->ks_form && !omega) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(h->device));
  int rc = set_batch(h, nsys, omega);
  if (rc) return rc;
  if ((rc = put_batch(h, h->b64, pr))) return rc;
  if ((rc = put_batch(h, h->x64, x0))) return rc;
  return 0;
}

int mtcocr_solve_resident(mtcocr_handle* h, double eps, int maxiter, int precision,
                          mtcocr_stats* stats) {
  if (!h || !stats || maxiter < 0 || !(eps > 0)) return MTCOCR_ERR_ARG;
  if (!h->have_matrix || !h->have_g || h->g_nb != h->nb || h->nsys < 1) return MTCOCR_ERR_STATE;
  if (precision != MTCOCR_PREC_FP64 && precision != MTCOCR_PREC_MIXED) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(h->device));
  int rc = with_sp(h->sp, [&](auto spc) -> int {
    constexpr int SP = decltype(spc)::value;
    if (precision == MTCOCR_PREC_FP64) return solve_fp64<SP>(h, eps, maxiter, stats);
    return solve_mixed<SP>(h, eps, maxiter, stats);
  });
  if (rc) return rc;
  for (int s = 0; s < h->nsys; ++s)
    if (stats[s].req != 0) return stats[s].req;
  return 0;
}

int mtcocr_get_solution(mtcocr_handle* h, double* x) {
  if (!h || !x) return MTCOCR_ERR_ARG;
  if (h->nsys < 1) return MTCOCR_ERR_STATE;
  HIPC(hipSetDevice(h->device));
  return get_batch(h, h->x64, x);
}

int mtcocr_solve_batch(mtcocr_handle* h, int nsys, const double* omega, const double* pr,
                       double* x, double eps, int maxiter, int precision, mtcocr_stats* stats) {
  if (!h || !pr || !x || !stats) return MTCOCR_ERR_ARG;
  auto t0 = std::chrono::steady_clock::now();
  int rc = mtcocr_set_rhs(h, nsys, omega, pr, x);
  if (rc) return rc;
  auto t1 = std::chrono::steady_clock::now();
  const int res = mtcocr_solve_resident(h, eps, maxiter, precision, stats);
  if (res < 0 && res <= MTCOCR_ERR_ARG && res != MTCOCR_ERR_NOTCONV) return res;
  auto t2 = std::chrono::steady_clock::now();
  rc = mtcocr_get_solution(h, x);
  if (rc) return rc;
  auto t3 = std::chrono::steady_clock::now();
  const double io = std::chrono::duration<double>(t1 - t0).count() +
                    std::chrono::duration<double>(t3 - t2).count();
  for (int s = 0; s < nsys; ++s) stats[s].seconds_io = io;
  return res;
}

int mtcocr_solve(mtcocr_handle* h, const double* pr, double* x, double eps, int maxiter,
                 int precision, mtcocr_stats* stats) {
  if (!h) return MTCOCR_ERR_ARG;
  if (h->ks_form) return MTCOCR_ERR_STATE;
  mtcocr_stats st;
  std::memset(&st, 0, sizeof(st));
  const int rc = mtcocr_solve_batch(h, 1, nullptr, pr, x, eps, maxiter, precision, &st);
  if (stats) *stats = st;
  return rc;
}

int mtcocr_pcocr_2x2_folded(int n, const int* ig, const int* jg, const int* idi, const int* ijg,
                            const double* di, const double* gg, const double* pr, double* x,
                            double eps, int maxiter, int kpar, int n_edges_c, int n_nodes_c,
                            const double* xyz, const int* nvkat, const int* nver,
                            const double* sigma3d, const int* edges, const int* ig_t,
                            const int* jg_t, const double* gg_t, const int* is_node_bound,
                            mtcocr_stats* stats) {
  (void)xyz;  // unused by the reference as well (only passed through)
  if (n <= 0 || n % 2 || n / 2 != n_edges_c) return MTCOCR_ERR_ARG;
  int dev = 0;
  if (const char* e = std::getenv("MTCOCR_DEVICE")) dev = std::atoi(e);
  mtcocr_handle* h = nullptr;
  int rc = mtcocr_create(dev, &h);
  if (rc) return rc;
  // locality order from the mesh when it is a structured hexahedral one (else the reference's)
  if (mtcocr_set_mesh_order(h, kpar, n_nodes_c, nver, n_edges_c, edges, MTCOCR_ORDER_GRID) !=
      MTCOCR_OK)
    (void)mtcocr_set_mesh_order(h, kpar, n_nodes_c, nver, n_edges_c, edges, 16);
  rc = mtcocr_set_matrix_block2x2(h, n / 2, ig, jg, idi, ijg, di, gg);
  if (!rc)
    rc = mtcocr_set_gradient_mesh(h, kpar, n_edges_c, n_nodes_c, nvkat, nver, sigma3d, edges,
                                  ig_t, jg_t, gg_t, is_node_bound);
  if (!rc) rc = mtcocr_solve(h, pr, x, eps, maxiter, MTCOCR_PREC_FP64, stats);
  mtcocr_destroy(h);
  return rc;
}

int mtcocr_bench_spmv(mtcocr_handle* h, int precision, int nsys, int reps, double* avg_ms,
                      double* bytes_per_launch) {
  if (!h || reps < 1 || !avg_ms || !bytes_per_launch) return MTCOCR_ERR_ARG;
  if (!h->have_matrix || h->nsys < 1 || nsys != h->nsys) return MTCOCR_ERR_STATE;
  HIPC(hipSetDevice(h->device));
  const int sp = h->sp;
  int rc = alloc_work(h, precision, sp);
  if (rc) return rc;
  hipStream_t st = h->stream;
  std::vector<SysScalars> hs;
  set_eps_status(h, hs, h->nsys, sp, 1e-8);
  HIPC(hipMemcpyAsync(h->scal.p, hs.data(), sizeof(SysScalars) * sp, hipMemcpyHostToDevice, st));
  int one = 1;
  HIPC(hipMemcpyAsync(h->nrun.p, &one, sizeof(int), hipMemcpyHostToDevice, st));
  const long long nv = (long long)h->nb * sp;
  const int gf = (int)std::min<long long>(2048, (nv + 255) / 256);
  float ms = 0.f;
  rc = with_sp(sp, [&](auto spc) -> int {
    constexpr int SP = decltype(spc)::value;
    if (precision == MTCOCR_PREC_FP64) {
      Engine<double, SP> E(h);
      k_fill<double2><<<gf, 256, 0, st>>>(E.s, nv);
      E.template spmv<0>(st, E.s, E.z, (const double2*)nullptr, (double2*)nullptr,
                         (double2*)nullptr);
      HIPC(hipEventRecord(h->ev0, st));
      for (int i = 0; i < reps; ++i)
        E.template spmv<0>(st, E.s, E.z, (const double2*)nullptr, (double2*)nullptr,
                           (double2*)nullptr);
      HIPC(hipEventRecord(h->ev1, st));
    } else {
      Engine<float, SP> E(h);
      k_fill<float2><<<gf, 256, 0, st>>>(E.s, nv);
      E.template spmv<0>(st, E.s, E.z, (const float2*)nullptr, (float2*)nullptr,
                         (float2*)nullptr);
      HIPC(hipEventRecord(h->ev0, st));
      for (int i = 0; i < reps; ++i)
        E.template spmv<0>(st, E.s, E.z, (const float2*)nullptr, (float2*)nullptr,
                           (float2*)nullptr);
      HIPC(hipEventRecord(h->ev1, st));
    }
    HIPC(hipGetLastError());
    HIPC(hipEventSynchronize(h->ev1));
    HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    return 0;
  });
  if (rc) return rc;
  *avg_ms = double(ms) / reps;
  // algorithmic bytes (independent of the layout): FP64 matrix values (16 B per complex,
  // 8 B per real non-zero) and int32 column indices once per launch, int32 row pointers, then
  // x read once and y written once per system (FP64 or FP32 vectors)
  const double ev = (precision == MTCOCR_PREC_FP64) ? 8.0 : 4.0;
  const double nnzc = double(h->nnz_c), nnzr = double(h->nnz_r);
  const double val = 2.0 * 8.0 * nnzc + 8.0 * nnzr;
  const double idx = 4.0 * (nnzc + nnzr);
  const double vec = double(nsys) * 2.0 * (2.0 * ev * h->nb);
  const double ptrs = 4.0 * (h->nb + 1.0);
  *bytes_per_launch = val + idx + vec + ptrs;
  return 0;
}

int mtcocr_bench_update(mtcocr_handle* h, int precision, int reps, double* avg_ms,
                        double* bytes_per_launch) {
  if (!h || reps < 1 || !avg_ms || !bytes_per_launch) return MTCOCR_ERR_ARG;
  if (precision != MTCOCR_PREC_FP64 && precision != MTCOCR_PREC_MIXED) return MTCOCR_ERR_ARG;
  if (!h->have_matrix || h->nsys < 1) return MTCOCR_ERR_STATE;
  HIPC(hipSetDevice(h->device));
  const int sp = h->sp;
  int rc = alloc_work(h, precision, sp);
  if (rc) return rc;
  hipStream_t st = h->stream;
  std::vector<SysScalars> hs;
  set_eps_status(h, hs, h->nsys, sp, 1e-8);
  for (auto& q : hs) q.mode = MRS_BLEND;  // the common step: the smoothing blend and the update
  HIPC(hipMemcpyAsync(h->scal.p, hs.data(), sizeof(SysScalars) * sp, hipMemcpyHostToDevice, st));
  int one = 1;
  HIPC(hipMemcpyAsync(h->nrun.p, &one, sizeof(int), hipMemcpyHostToDevice, st));
  const long long nv = (long long)h->nb * sp;
  const int gf = (int)std::min<long long>(2048, (nv + 255) / 256);
  float ms = 0.f;
  int streams = 0;
  double node_bytes = 0.0;  // help gathered once per node and system, the diagonal once per edge
  rc = with_sp(sp, [&](auto spc) -> int {
    constexpr int SP = decltype(spc)::value;
    auto run = [&](auto tag) -> int {
      using T = decltype(tag);
      using T2 = typename Vec2<T>::type;
      Engine<T, SP> E(h);
      for (T2* v : {E.x, E.x2, E.p, E.r, E.a, E.s, E.w, E.ym, E.sm, E.sr[1]})
        k_fill<T2><<<gf, 256, 0, st>>>(v, nv);
      k_fill<T2><<<gf, 256, 0, st>>>(E.help, (long long)std::max(E.m, 1) * SP);
      T2* ym = E.ym_or_null();
      // the pass as the solver runs it: FP64 without x and p (8 vectors: r, sm, s, a, w read;
      // r, sm and the next s written; after the fused grid preconditioner w is formed from the
      // node vector help instead of read), FP32 vectors with x, p and the smoothed iterate
      bool formed = false;
      if constexpr (Engine<T, SP>::kLazy && pg_width(SP)) formed = E.use_pg;
      streams = Engine<T, SP>::kLazy ? (formed ? 7 : 8) : ym ? 14 : 12;
      node_bytes = formed ? double(E.m) * 16.0 * double(h->nsys) + 16.0 * double(h->nb) : 0.0;
      auto launch = [&](int i) {
        if constexpr (Engine<T, SP>::kLazy && pg_width(SP)) {
          if (formed) {
            k_xr_fused<T, SP, true, true><<<E.grid_u, kBlock, 0, st>>>(
                nullptr, nullptr, nullptr, E.r, E.a, i & 1 ? E.sr[1] : E.s,
                i & 1 ? E.s : E.sr[1], E.help, nullptr, E.sm, E.sc, E.nrun, E.part, E.nb,
                E.wrec());
            return;
          }
        }
        if constexpr (Engine<T, SP>::kLazy)
          k_xr_fused<T, SP, true><<<E.grid_u, kBlock, 0, st>>>(
              nullptr, nullptr, nullptr, E.r, E.a, i & 1 ? E.sr[1] : E.s, i & 1 ? E.s : E.sr[1],
              E.w, nullptr, E.sm, E.sc, E.nrun, E.part, E.nb);
        else
          k_xr_fused<T, SP><<<E.grid_u, kBlock, 0, st>>>(i & 1 ? E.x2 : E.x, i & 1 ? E.x : E.x2,
                                                         E.p, E.r, E.a, E.s, E.s, E.w, ym, E.sm,
                                                         E.sc, E.nrun, E.part, E.nb);
      };
      launch(1);
      HIPC(hipEventRecord(h->ev0, st));
      for (int i = 0; i < reps; ++i) launch(i);
      HIPC(hipEventRecord(h->ev1, st));
      return 0;
    };
    const int r = (precision == MTCOCR_PREC_FP64) ? run(double()) : run(float());
    if (r) return r;
    HIPC(hipGetLastError());
    HIPC(hipEventSynchronize(h->ev1));
    HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    return 0;
  });
  if (rc) return rc;
  *avg_ms = double(ms) / reps;
  // algorithmic bytes: each stream read or written once per system and edge (plus, when w is
  // formed in the pass, help once per interior node and system and the diagonal once per edge)
  const double ev = (precision == MTCOCR_PREC_FP64) ? 16.0 : 8.0;
  *bytes_per_launch = double(streams) * ev * double(h->nb) * double(h->nsys) + node_bytes;
  return 0;
}

int mtcocr_bench_flush(mtcocr_handle* h, int reps, double* avg_ms, double* bytes_per_launch) {
  if (!h || reps < 1 || !avg_ms || !bytes_per_launch) return MTCOCR_ERR_ARG;
  if (!h->have_matrix || h->nsys < 1) return MTCOCR_ERR_STATE;
  HIPC(hipSetDevice(h->device));
  const int sp = h->sp;
  int rc = alloc_work(h, MTCOCR_PREC_FP64, sp);
  if (rc) return rc;
  hipStream_t st = h->stream;
  std::vector<SysScalars> hs;
  set_eps_status(h, hs, h->nsys, sp, 1e-8);  // running systems: every lane takes all eight steps
  HIPC(hipMemcpyAsync(h->scal.p, hs.data(), sizeof(SysScalars) * sp, hipMemcpyHostToDevice, st));
  const int xp = h->xp_period;
  const int kn[2] = {1, xp};  // [0] running count, [1] the iteration a flush ends
  HIPC(hipMemcpyAsync(h->nrun.p, kn, sizeof(kn), hipMemcpyHostToDevice, st));
  const long long nv = (long long)h->nb * sp;
  const int gf = (int)std::min<long long>(2048, (nv + 255) / 256);
  float ms = 0.f;
  rc = with_sp(sp, [&](auto spc) -> int {
    constexpr int SP = decltype(spc)::value;
    Engine<double, SP> E(h);
    XpRing rs{};
    for (int i = 0; i < xp; ++i) rs.s[i] = E.sr[i];
    k_fill<double2><<<gf, 256, 0, st>>>(E.x, nv);
    k_fill<double2><<<gf, 256, 0, st>>>(E.p, nv);
    for (int i = 0; i < xp; ++i) k_fill<double2><<<gf, 256, 0, st>>>(E.sr[i], nv);
    auto launch = [&]() {
      if (xp == 16)
        k_flush_xp<SP, 16><<<E.grid_u, kBlock, 0, st>>>(E.x, E.p, rs, E.sc, E.nrun, E.nb);
      else
        k_flush_xp<SP, 8><<<E.grid_u, kBlock, 0, st>>>(E.x, E.p, rs, E.sc, E.nrun, E.nb);
    };
    launch();
    HIPC(hipEventRecord(h->ev0, st));
    for (int i = 0; i < reps; ++i) launch();
    HIPC(hipEventRecord(h->ev1, st));
    HIPC(hipGetLastError());
    HIPC(hipEventSynchronize(h->ev1));
    HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    return 0;
  });
  if (rc) return rc;
  *avg_ms = double(ms) / reps;
  // x, p and eight s read, x and p written, once per system and edge
  *bytes_per_launch = double(xp + 4) * 16.0 * double(h->nb) * double(h->nsys);
  return 0;
}

int mtcocr_bench_copy(mtcocr_handle* h, int64_t nbytes, int reps, double* avg_ms) {
  if (!h || nbytes < 16 || nbytes % 16 || reps < 1 || !avg_ms) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(h->device));
  DevBuf src, dst;
  int rc;
  if ((rc = ensure(h, src, (size_t)nbytes)) || (rc = ensure(h, dst, (size_t)nbytes))) return rc;
  hipStream_t st = h->stream;
  const long long n = nbytes / 16;
  k_fill<double2><<<2048, 256, 0, st>>>(src.as<double2>(), n);
  k_copy<<<8192, 256, 0, st>>>(src.as<double2>(), dst.as<double2>(), n);
  HIPC(hipEventRecord(h->ev0, st));
  for (int i = 0; i < reps; ++i)
    k_copy<<<8192, 256, 0, st>>>(src.as<double2>(), dst.as<double2>(), n);
  HIPC(hipEventRecord(h->ev1, st));
  HIPC(hipGetLastError());
  HIPC(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *avg_ms = double(ms) / reps;
  return 0;
}

int mtcocr_set_launch(mtcocr_handle* h, int cap_spmv, int cap_vec, int cap_node,
                      int spmv_layout) {
  if (!h) return MTCOCR_ERR_ARG;
  auto bad = [](int c) { return c != 0 && (c < kXcd || c > kMaxBlocks || c % kXcd); };
  if (bad(cap_spmv) || bad(cap_vec) || bad(cap_node)) return MTCOCR_ERR_ARG;
  if (spmv_layout < -1 || spmv_layout > 4) return MTCOCR_ERR_ARG;
  h->set_caps(cap_spmv, cap_vec, cap_node, spmv_layout);
  drop_graph(h);  // the captured iterations hold the old grids
  return 0;
}

}  // extern "C"

namespace {
// the SpMV layout that runs for precision q at batch-width index k (row blocks fall back to
// the row pack where they do not apply)
int effective_layout(const mtcocr_handle* h, int q, int k) {
  int l = h->spmv_layout[q][k];
  if (l == 3 && h->stc_big) l = 4;  // the wave-task kernel cannot address the table
  if (l >= 3 && !(h->stc_ok && (k == 4 || (k == 3 && l == 4)))) l = 2;
  if (l == 2 && (q == 0 ? blk_lds_bytes<double>(h) : blk_lds_bytes<float>(h)) == 0) l = 1;
  if (!rows_fit(h, 1 << k, q == 0 ? 16 : 8)) l = 0;
  return l;
}
}  // namespace

extern "C" {

int mtcocr_set_stencil_sweep(mtcocr_handle* h, int kc) {
  if (!h || kc < 1) return MTCOCR_ERR_ARG;
  for (auto& row : h->stc_kc)
    for (int& c : row) c = kc;
  drop_graph(h);  // captured launches hold the old task count
  return 0;
}

int mtcocr_set_stencil_band(mtcocr_handle* h, int band) {
  if (!h || band < 0) return MTCOCR_ERR_ARG;
  for (auto& row : h->stl_band)
    for (int& c : row) c = band;
  drop_graph(h);  // captured launches hold the old task order
  return 0;
}

int mtcocr_set_xp_period(mtcocr_handle* h, int period) {
  if (!h || (period != 0 && period != 8 && period != 16)) return MTCOCR_ERR_ARG;
  h->xp_request = period;
  if (period == 8) {
    if (h->xp_period != 8) drop_graph(h);  // captured chunks name the buffers freed below
    for (int i = 15; i < kRingMax - 1; ++i) h->vsr[i].release();
    h->xp_period = 8;
  }
  return 0;
}

int mtcocr_get_xp_period(mtcocr_handle* h, int* period) {
  if (!h || !period) return MTCOCR_ERR_ARG;
  *period = h->xp_period;
  return 0;
}

int mtcocr_get_stencil_band(mtcocr_handle* h, int precision, int* band) {
  if (!h || !band) return MTCOCR_ERR_ARG;
  if (precision != MTCOCR_PREC_FP64 && precision != MTCOCR_PREC_MIXED) return MTCOCR_ERR_ARG;
  *band = h->stl_band[precision][mtcocr_handle::sp_index(h->nsys > 0 ? h->sp : 1)];
  return 0;
}

int mtcocr_get_stencil_sweep(mtcocr_handle* h, int precision, int* kc) {
  if (!h || !kc) return MTCOCR_ERR_ARG;
  if (precision != MTCOCR_PREC_FP64 && precision != MTCOCR_PREC_MIXED) return MTCOCR_ERR_ARG;
  *kc = h->stc_kc[precision][mtcocr_handle::sp_index(h->nsys > 0 ? h->sp : 1)];
  if (h->grid_order && h->grid.nz > 0) *kc = std::min(*kc, h->grid.nz);  // whole columns
  return 0;
}

int mtcocr_get_launch(mtcocr_handle* h, int precision, int* cap_spmv, int* cap_vec,
                      int* cap_node, int* spmv_layout) {
  if (!h) return MTCOCR_ERR_ARG;
  if (precision != MTCOCR_PREC_FP64 && precision != MTCOCR_PREC_MIXED) return MTCOCR_ERR_ARG;
  const int k = mtcocr_handle::sp_index(h->nsys > 0 ? h->sp : 1), q = precision;
  if (cap_spmv) *cap_spmv = h->cap_spmv[q][k];
  if (cap_vec) *cap_vec = h->cap_vec[q][k];
  if (cap_node) *cap_node = h->cap_node[q][k];
  if (spmv_layout) *spmv_layout = effective_layout(h, q, k);
  return 0;
}

// Launch-shape autotuner. For the batch currently on the handle (its systems, padding and
// frequencies) it times the kernel families whose launch is tunable, each at every candidate
// workgroup cap, on zero-filled work vectors (the kernels' memory traffic does not depend on
// the values), and keeps the fastest choice of each family:
//   SpMV            layout (sliced ELL k_spmv, row pack k_spmv_rg for 4..64 systems, row
//                   blocks k_spmv_blk for 4..16 systems) x cap
//                   (the cap also sets the partial count of its reductions)
//   node pass       k_pc_node cap
//   edge streams    k_pc_edge_iter cap (reported as cap_vec), timed after the node pass at
//                   its tuned cap, as it runs in an iteration
//   update pass     k_xr_fused, tuned apart (set_launch's cap_vec sets both)
// An FP64 batch is tuned at its own width and then at every width its compaction narrows it
// to (sp/2, sp/4, ... 1, on its first systems), so that the narrowed batches also run at tuned
// shapes. Only work buffers are written; the resident right-hand sides and solutions are
// untouched.
int mtcocr_autotune(mtcocr_handle* h, int precision, int reps, int* cap_spmv, int* cap_vec,
                    int* cap_node, int* spmv_layout, double* spmv_ms) {
  if (!h || reps < 1) return MTCOCR_ERR_ARG;
  if (precision != MTCOCR_PREC_FP64 && precision != MTCOCR_PREC_MIXED) return MTCOCR_ERR_ARG;
  if (!h->have_matrix || !h->have_g || h->g_nb != h->nb || h->nsys < 1) return MTCOCR_ERR_STATE;
  HIPC(hipSetDevice(h->device));
  // FP64 batches wider than kGroupWidth iterate in groups of kGroupWidth: tune that width
  const int sp =
      (precision == MTCOCR_PREC_FP64 && h->sp > kGroupWidth) ? kGroupWidth : h->sp;
  // the batch's own width, then (FP64) every width its compaction narrows it to: sp/2, sp/4, ... 1
  double ms0 = 0.0;
  int rc = autotune_width(h, precision, reps, sp, &ms0);
  if (rc) return rc;
  if (precision == MTCOCR_PREC_FP64) {
    for (int w = sp / 2; w >= 1; w /= 2) {
      double msw = 0.0;
      if ((rc = autotune_width(h, precision, reps, w, &msw))) return rc;
    }
  }
  const int ki = mtcocr_handle::sp_index(sp), q = precision;
  if (cap_spmv) *cap_spmv = h->cap_spmv[q][ki];
  if (cap_node) *cap_node = h->cap_node[q][ki];
  if (cap_vec) *cap_vec = h->cap_vec[q][ki];
  if (spmv_layout) *spmv_layout = effective_layout(h, q, ki);
  if (spmv_ms) *spmv_ms = ms0;
  return 0;
}

int mtcocr_apply(mtcocr_handle* h, int op, int precision, const double* x, double* y) {
  if (!h || !x || !y || (op != 0 && op != 1)) return MTCOCR_ERR_ARG;
  if (precision != MTCOCR_PREC_FP64 && precision != MTCOCR_PREC_MIXED) return MTCOCR_ERR_ARG;
  if (!h->have_matrix || (op == 1 && (!h->have_g || h->g_nb != h->nb))) return MTCOCR_ERR_STATE;
  HIPC(hipSetDevice(h->device));
  int rc;
  if (h->nsys < 1) {
    if (h->ks_form) return MTCOCR_ERR_STATE;
    if ((rc = set_batch(h, 1, nullptr))) return rc;
  }
  const int sp = h->sp;
  if ((rc = alloc_work(h, precision, sp))) return rc;
  hipStream_t st = h->stream;
  std::vector<SysScalars> hs;
  set_eps_status(h, hs, 1, sp, 1e-8);  // system 0 only
  HIPC(hipMemcpyAsync(h->scal.p, hs.data(), sizeof(SysScalars) * sp, hipMemcpyHostToDevice, st));
  int one = 1;
  HIPC(hipMemcpyAsync(h->nrun.p, &one, sizeof(int), hipMemcpyHostToDevice, st));
  // x -> r64 (system 0 of the [row][system] layout), converted on the host for FP32
  const int nb = h->nb;
  std::vector<double2> in((size_t)nb * sp);
  for (int r = 0; r < nb; ++r)
    for (int s = 0; s < sp; ++s) {
      const size_t o = h->order.empty() ? (size_t)r : (size_t)h->order[r];
      in[(size_t)r * sp + s] = s == 0 ? make_double2(x[2 * o], x[2 * o + 1]) : make_double2(0, 0);
    }
  std::vector<double2> out((size_t)nb * sp);
  rc = with_sp(sp, [&](auto spc) -> int {
    constexpr int SP = decltype(spc)::value;
    auto run = [&](auto tag) -> int {
      using T = decltype(tag);
      using T2 = typename Vec2<T>::type;
      Engine<T, SP> E(h);
      std::vector<T2> xin((size_t)nb * SP), yout((size_t)nb * SP);
      for (size_t k = 0; k < xin.size(); ++k) {
        xin[k].x = T(in[k].x);
        xin[k].y = T(in[k].y);
      }
      HIPC(hipMemcpyAsync(E.r, xin.data(), xin.size() * sizeof(T2), hipMemcpyHostToDevice, st));
      if (op == 0) {
        E.template spmv<0>(st, E.r, E.z, (const T2*)nullptr, (T2*)nullptr, (T2*)nullptr);
      } else {
        E.setup_diagonals(st);
        k_pc_node<T, SP, false><<<E.grid_n, kBlock, 0, st>>>(E.G, E.r, nullptr, E.dn, E.help,
                                                             E.sc, E.nrun);
        k_pc_edge_plain<T, SP><<<E.grid_e, kBlock, 0, st>>>(E.G, E.r, E.da, E.help, E.z, E.sc,
                                                            E.nrun);
      }
      HIPC(hipGetLastError());
      HIPC(hipMemcpyAsync(yout.data(), E.z, yout.size() * sizeof(T2), hipMemcpyDeviceToHost, st));
      HIPC(hipStreamSynchronize(st));
      for (size_t k = 0; k < yout.size(); ++k) out[k] = make_double2(double(yout[k].x), double(yout[k].y));
      return 0;
    };
    if (precision == MTCOCR_PREC_FP64) return run(double());
    return run(float());
  });
  if (rc) return rc;
  for (int r = 0; r < nb; ++r) {
    const size_t o = h->order.empty() ? (size_t)r : (size_t)h->order[r];
    y[2 * o] = out[(size_t)r * sp].x;
    y[2 * o + 1] = out[(size_t)r * sp].y;
  }
  return 0;
}

int mtcocr_apply_noalloc(mtcocr_handle* h, int op, double omega, const double* x, double* y,
                         void* stream) {
  if (!h || !x || !y || (op != 0 && op != 1)) return MTCOCR_ERR_ARG;
  if (!h->have_matrix || (op == 1 && (!h->have_g || h->g_nb != h->nb))) return MTCOCR_ERR_STATE;
  if (!h->order.empty()) return MTCOCR_ERR_STATE;  // x and y are in the caller's row order
  if (!h->ks_form && omega != 1.0) return MTCOCR_ERR_ARG;
  if (!(omega > 0.0)) return MTCOCR_ERR_ARG;
  HIPC(hipSetDevice(h->device));
  hipError_t e = hipSuccess;
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr(h->device, &e));
  if (!base) {
    std::snprintf(g_err, sizeof(g_err), "scratch address lookup failed: %s", hipGetErrorString(e));
    return MTCOCR_ERR_HIP;
  }
  // scratch layout (256-byte aligned parts): scalars | running count, omega | SpMV partials
  // (one (re, im) pair per workgroup) | D_N^-1 | help (m complex values each)
  const int grid_s = grid_for(h->nb, h->cap_spmv[0][0]);
  const int grid_e = grid_for(h->nb, h->cap_vec[0][0]);
  const int grid_n = grid_for(std::max(h->m, 1), h->cap_node[0][0]);
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t o_sc = 0, o_run = al(sizeof(SysScalars)), o_part = o_run + 256;
  const size_t o_dn = o_part + al(sizeof(double) * 2 * (size_t)grid_s);
  const size_t o_help = o_dn + al(sizeof(double2) * (size_t)std::max(h->m, 1));
  const size_t need = (op == 0) ? o_dn : o_help + al(sizeof(double2) * (size_t)std::max(h->m, 1));
  if (need > kScratchBytes) {
    std::snprintf(g_err, sizeof(g_err),
                  "mtcocr_apply_noalloc needs %zu bytes of scratch, the limit is %zu (m = %d)",
                  need, kScratchBytes, h->m);
    return MTCOCR_ERR_ARG;
  }
  SysScalars* sc = reinterpret_cast<SysScalars*>(base + o_sc);
  int* nrun = reinterpret_cast<int*>(base + o_run);
  double* om = reinterpret_cast<double*>(base + o_run + 64);
  double* part = reinterpret_cast<double*>(base + o_part);
  double2* dn = reinterpret_cast<double2*>(base + o_dn);
  double2* help = reinterpret_cast<double2*>(base + o_help);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : h->stream;
  const double2* xv = reinterpret_cast<const double2*>(x);
  double2* yv = reinterpret_cast<double2*>(y);
  k_noalloc_init<<<1, 64, 0, st>>>(sc, nrun, om, omega);
  if (op == 0) {
    // the sliced-ELL SpMV at one system, as mtcocr_apply runs it without a batch
    launch_spmv<double, 1, 0>(h, 0, grid_s, st, om, xv, yv, nullptr, nullptr, nullptr, sc, nrun,
                              part);
  } else {
    const int g2 = std::min(4096, (int)(((long long)h->m + 255) / 256) + 1);
    k_diag_n<double, 1><<<g2, 256, 0, st>>>(ell_view<double>(h), g_view<double>(h), om, dn);
    const GView<double> G = g_view<double>(h);
    k_pc_node<double, 1, false><<<grid_n, kBlock, 0, st>>>(G, xv, nullptr, dn, help, sc, nrun);
    DiagA da;
    da.a = h->da64.as<double>();
    da.b = h->db64.as<double>();
    da.omega = om;
    k_pc_edge_plain<double, 1><<<grid_e, kBlock, 0, st>>>(G, xv, da, help, yv, sc, nrun);
  }
  HIPC(hipGetLastError());
  return 0;
}

int mtcocr_matrix_info(mtcocr_handle* h, int64_t* nnz_complex, int64_t* nnz_real, int64_t* m,
                       int64_t* device_bytes) {
  if (!h) return MTCOCR_ERR_ARG;
  if (nnz_complex) *nnz_complex = h->nnz_c;
  if (nnz_real) *nnz_real = h->nnz_r;
  if (m) *m = h->m;
  if (device_bytes) *device_bytes = h->device_bytes();
  return 0;
}

int mtcocr_sigma_gradient(mtcocr_handle* h, int nxn, int nyn, int nzn, const double* x,
                          const double* y, const double* z, int nsys, const double* coef,
                          const double* u, const double* v, double* grad) {
  if (!h || !x || !y || !z || !coef || !u || !v || !grad || nsys < 1) return MTCOCR_ERR_ARG;
  if (nxn < 2 || nyn < 2 || nzn < 2) return MTCOCR_ERR_ARG;
  const long long ncx = nxn - 1, ncy = nyn - 1, ncz = nzn - 1, ncells = ncx * ncy * ncz;
  const long long nb = (long long)mtmodel_num_edges(nxn, nyn, nzn);
  if (nb <= 0 || nb > 2147483000LL) return MTCOCR_ERR_ARG;
  // the cells' edges in the mesh writer's local order (edges.dat rows, 1-based)
  std::vector<int> ed((size_t)ncells * 25), ce((size_t)ncells * 12);
  int64_t kt1 = 0;
  if (mtmodel_topology(nxn, nyn, nzn, nullptr, ed.data(), nullptr, nullptr, &kt1) != 0)
    return MTCOCR_ERR_ARG;
  for (long long c = 0; c < ncells; ++c)
    for (int k = 0; k < 12; ++k) ce[(size_t)c * 12 + k] = ed[(size_t)c * 25 + k] - 1;
  std::vector<double> xyz;
  xyz.reserve((size_t)nxn + nyn + nzn);
  xyz.insert(xyz.end(), x, x + nxn);
  xyz.insert(xyz.end(), y, y + nyn);
  xyz.insert(xyz.end(), z, z + nzn);
  HIPC(hipSetDevice(h->device));
  hipStream_t st = h->stream;
  const size_t vbytes = (size_t)nsys * nb * 2 * sizeof(double);
  auto put = [&](DevBuf& d, const void* src, size_t bytes) -> int {
    if (d.bytes < bytes || !d.p) HIPC(d.alloc(bytes));
    if (src) HIPC(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, st));
    return 0;
  };
  int rc;
  if ((rc = put(h->sg_ce, ce.data(), ce.size() * sizeof(int))) ||
      (rc = put(h->sg_xyz, xyz.data(), xyz.size() * sizeof(double))) ||
      (rc = put(h->sg_coef, coef, (size_t)nsys * 2 * sizeof(double))) ||
      (rc = put(h->sg_u, u, vbytes)) || (rc = put(h->sg_v, v, vbytes)) ||
      (rc = put(h->sg_out, nullptr, (size_t)ncells * sizeof(double))))
    return rc;
  const double* xd = h->sg_xyz.as<double>();
  const int grid = (int)std::min<long long>((ncells + 255) / 256, 65536);
  k_sigma_grad<<<grid, 256, 0, st>>>(h->sg_ce.as<int>(), xd, xd + nxn, xd + nxn + nyn, int(ncx),
                                     int(ncy), ncells, nsys, h->sg_coef.as<double>(),
                                     h->sg_u.as<double2>(), h->sg_v.as<double2>(), nb,
                                     h->sg_out.as<double>());
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(grad, h->sg_out.p, (size_t)ncells * sizeof(double), hipMemcpyDeviceToHost,
                      st));
  HIPC(hipStreamSynchronize(st));
  return 0;
}

}  // extern "C"
