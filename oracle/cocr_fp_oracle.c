/*
 * ORACLE — test infrastructure only (see cocr_fp_oracle.h for the rules and the pinning list).
 *
 * FP64 CPU restatement of the reference's Code/COCR_FP, operation for operation:
 *   RCI state machine       Code/COCR_FP/pcocr_rci.cpp:50-181, rci.cpp:24-90
 *   driver / request loop   Code/COCR_FP/pcocr.cpp:40-126
 *   SpMV (lower-tri blocks) Code/COCR_FP/block_2x2_solver.cpp:41-53,73-173
 *   folded preconditioner   Code/COCR_FP/FoldedPreconditioner.cpp:68-259,473-736
 *   MRS (real variant)      Code/COCR_FP/MRS.cpp:27-39,85-153,244-248
 *   vector kernels          Code/COCR_FP/base_solver.cpp:41-69,382-399,433-452
 *   kit file                Code/COCR_FP/base_solver.cpp:181-200
 *   process / file contract Code/COCR_FP/HarmVect3d.cpp:36-365, vec_prep_data.cpp:55-187,
 *                           in_out.cpp:212-304,492-526, ControlOMP.cpp:41-60
 * With one thread every loop runs in the reference's serial order. With T threads the loops are
 * split into the contiguous chunks of an OpenMP static schedule and per-thread partial sums are
 * combined in thread order (deterministic for a fixed T).
 */
#include "cocr_fp_oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------------------------ */
/* helpers                                                                                    */

static void chunk_of(long n, int T, int t, long* lo, long* hi) {
  long q = n / T, r = n % T;
  if (t < r) {
    *lo = (long)t * (q + 1);
    *hi = *lo + q + 1;
  } else {
    *lo = (long)t * q + r;
    *hi = *lo + q;
  }
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

#define MAXT 256

/* Base_solver::Scal (base_solver.cpp:41-56) */
static double scal(const double* x, const double* y, long n, int T) {
  double part[MAXT];
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    long lo, hi, i;
    double s = 0.0;
    chunk_of(n, T, t, &lo, &hi);
    for (i = lo; i < hi; i++) s += x[i] * y[i];
    part[t] = s;
  }
  double sum = 0.0;
  for (int t = 0; t < T; t++) sum += part[t];
  return sum;
}

/* Base_solver::Norm_Euclid (base_solver.cpp:60-69) */
static double norm_euclid(const double* x, long n, int T) {
  double t = scal(x, x, n, T);
  if (t > 0.0) return sqrt(t);
  return 0.0;
}

/* Base_solver::ScalCmplx (base_solver.cpp:382-399): sum x_i*y_i without conjugation */
void orc_scal_cmplx(const double* x, const double* y, int nb, int T, double* re, double* im) {
  double pre[MAXT], pim[MAXT];
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    long lo, hi, i;
    double sr = 0.0, si = 0.0;
    chunk_of(nb, T, t, &lo, &hi);
    for (i = lo; i < hi; i++) {
      sr += x[i * 2] * y[i * 2] - x[i * 2 + 1] * y[i * 2 + 1];
      si += x[i * 2] * y[i * 2 + 1] + x[i * 2 + 1] * y[i * 2];
    }
    pre[t] = sr;
    pim[t] = si;
  }
  double sr = 0.0, si = 0.0;
  for (int t = 0; t < T; t++) {
    sr += pre[t];
    si += pim[t];
  }
  *re = sr;
  *im = si;
}

/* Base_solver::Cmplx_axpy (base_solver.cpp:433-452): z = x + a*y */
static void cmplx_axpy(double a_re, double a_im, const double* x, const double* y, double* z,
                       long nb, int T) {
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    long lo, hi, i;
    chunk_of(nb, T, t, &lo, &hi);
    for (i = lo; i < hi; i++) {
      double t_re = x[i * 2] + (y[i * 2] * a_re - y[i * 2 + 1] * a_im);
      double t_im = x[i * 2 + 1] + (y[i * 2] * a_im + y[i * 2 + 1] * a_re);
      z[i * 2] = t_re;
      z[i * 2 + 1] = t_im;
    }
  }
}

/* complex division with the C99 semantics used by std::complex<double> in the reference */
static void cdiv(double ar, double ai, double br, double bi, double* qr, double* qi) {
  double complex a = ar + ai * I;
  double complex b = br + bi * I;
  double complex q = a / b;
  *qr = creal(q);
  *qi = cimag(q);
}

/* ------------------------------------------------------------------------------------------ */
/* SpMV: Block_2x2_solver::Mult_block_2x2 + Mult_MV_block_2x2 (block_2x2_solver.cpp:41-173)   */

static inline void mult_block(const double* a, int size, const double* x, double* y) {
  if (size == 2) {
    y[0] += a[0] * x[0] - a[1] * x[1];
    y[1] += a[1] * x[0] + a[0] * x[1];
  } else {
    y[0] += a[0] * x[0];
    y[1] += a[0] * x[1];
  }
}

void orc_spmv(int nb, const int* ig, const int* jg, const int* idi, const int* ijg,
              const double* di, const double* gg, const double* x, double* y, int T,
              double* y_omp) {
  long i, j;
  if (T == 1 || nb < 10) {
    for (i = 0; i < (long)nb * 2; i++) y[i] = 0.0;
    for (i = 0; i < nb; i++) {
      long ib = i * 2;
      mult_block(&di[idi[i]], idi[i + 1] - idi[i], &x[ib], &y[ib]);
      for (j = ig[i]; j <= ig[i + 1] - 1; j++) {
        long jb = (long)jg[j] * 2;
        int k = ijg[j];
        int size = ijg[j + 1] - k;
        mult_block(&gg[k], size, &x[jb], &y[ib]);
        mult_block(&gg[k], size, &x[ib], &y[jb]);
      }
    }
    return;
  }
  /* threads > 1: thread 0 scatters into y, thread t>0 into its private copy, then the copies
     are added in thread order */
  int own = (y_omp == NULL);
  if (own) y_omp = (double*)malloc(sizeof(double) * 2 * (size_t)nb * (size_t)(T - 1));
  for (i = 0; i < (long)nb * 2; i++) y[i] = 0.0;
  for (i = 0; i < (long)nb * 2 * (T - 1); i++) y_omp[i] = 0.0;
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    long lo, hi, r, q;
    double* yy = (t == 0) ? y : y_omp + (size_t)(t - 1) * 2 * (size_t)nb;
    chunk_of(nb, T, t, &lo, &hi);
    for (r = lo; r < hi; r++) {
      long ib = r * 2;
      mult_block(&di[idi[r]], idi[r + 1] - idi[r], &x[ib], &yy[ib]);
      for (q = ig[r]; q <= ig[r + 1] - 1; q++) {
        long jb = (long)jg[q] * 2;
        int k = ijg[q];
        int size = ijg[q + 1] - k;
        mult_block(&gg[k], size, &x[jb], &yy[ib]);
        mult_block(&gg[k], size, &x[ib], &yy[jb]);
      }
    }
  }
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    long lo, hi, r;
    chunk_of((long)nb * 2, T, t, &lo, &hi);
    for (r = lo; r < hi; r++)
      for (int k = 0; k < T - 1; k++) y[r] += y_omp[(size_t)k * 2 * (size_t)nb + r];
  }
  if (own) free(y_omp);
}

/* ------------------------------------------------------------------------------------------ */
/* Folded preconditioner                                                                      */

void orc_dfp_free(orc_dfp* p) {
  free(p->iptr_g);
  free(p->jptr_g);
  free(p->gg_g);
  free(p->iptr_gt);
  free(p->jptr_gt);
  free(p->gg_gt);
  free(p->diag_a);
  free(p->diag_gagt);
  free(p->help);
  memset(p, 0, sizeof(*p));
}

/* FoldedPreconditioner::BuildGMatrix (FoldedPreconditioner.cpp:68-259) */
int orc_build_g(orc_dfp* p, int kpar, int n_edges_c, int n_nodes_c, const int* nvkat,
                const int* nver, const double* sigma3d, const int* edges, const int* ig_t,
                const int* jg_t, const double* gg_t, const int* is_node_bound) {
  int i, j, k;
  free(p->iptr_g);
  free(p->jptr_g);
  free(p->gg_g);
  free(p->iptr_gt);
  free(p->jptr_gt);
  free(p->gg_gt);
  p->n = n_edges_c;
  int* in_sigma = (int*)calloc((size_t)n_nodes_c + 1, sizeof(int));
  int* renum = (int*)malloc(sizeof(int) * ((size_t)n_nodes_c + 1));
  if (!in_sigma || !renum) return -12;
  for (i = 0; i < kpar; i++) {
    if (sigma3d[nvkat[i]] > 0.0) {
      for (j = 0; j < 8; j++) {
        int node = nver[(size_t)i * 14 + j];
        if (node < n_nodes_c) {
          if (is_node_bound[node]) continue;
          in_sigma[node] = 1;
        }
      }
    }
  }
  k = 0;
  for (i = 0; i < n_nodes_c; i++) {
    renum[i] = -1;
    if (in_sigma[i]) renum[i] = k++;
  }
  p->m = k;
  p->iptr_gt = (int*)malloc(sizeof(int) * ((size_t)n_edges_c + 1));
  p->iptr_gt[0] = 0;
  for (i = 0; i < n_edges_c; i++) {
    j = 0;
    for (k = 0; k < 2; k++) {
      int v = edges[(size_t)i * 2 + k];
      if (v < n_nodes_c) {
        if (in_sigma[v]) j++;
      } else {
        for (int ii = ig_t[v]; ii < ig_t[v + 1]; ii++)
          if (in_sigma[jg_t[ii]]) j++;
      }
    }
    p->iptr_gt[i + 1] = p->iptr_gt[i] + j;
  }
  int nnz = p->iptr_gt[n_edges_c];
  p->jptr_gt = (int*)malloc(sizeof(int) * ((size_t)nnz + 1));
  p->gg_gt = (double*)malloc(sizeof(double) * ((size_t)nnz + 1));
  j = 0;
  for (i = 0; i < n_edges_c; i++) {
    for (k = 0; k < 2; k++) {
      int v = edges[(size_t)i * 2 + k];
      if (v < n_nodes_c) {
        if (in_sigma[v]) {
          p->jptr_gt[j] = renum[v];
          p->gg_gt[j] = (k == 0) ? -0.5 : 0.5;
          j++;
        }
      } else {
        for (int ii = ig_t[v]; ii < ig_t[v + 1]; ii++) {
          int kk = jg_t[ii];
          if (in_sigma[kk]) {
            p->jptr_gt[j] = renum[kk];
            p->gg_gt[j] = (k == 0) ? -0.5 * gg_t[ii] : 0.5 * gg_t[ii];
            j++;
          }
        }
      }
    }
  }
  /* transpose into G (node rows, edges ascending) */
  int* temp = (int*)calloc((size_t)p->m + 1, sizeof(int));
  for (i = 0; i < n_edges_c; i++)
    for (j = p->iptr_gt[i]; j < p->iptr_gt[i + 1]; j++) temp[p->jptr_gt[j]]++;
  p->iptr_g = (int*)malloc(sizeof(int) * ((size_t)p->m + 1));
  p->jptr_g = (int*)malloc(sizeof(int) * ((size_t)nnz + 1));
  p->gg_g = (double*)malloc(sizeof(double) * ((size_t)nnz + 1));
  p->iptr_g[0] = 0;
  for (i = 0; i < p->m; i++) p->iptr_g[i + 1] = p->iptr_g[i] + temp[i];
  for (i = 0; i < p->m; i++) temp[i] = p->iptr_g[i];
  for (i = 0; i < n_edges_c; i++)
    for (j = p->iptr_gt[i]; j < p->iptr_gt[i + 1]; j++) {
      k = p->jptr_gt[j];
      p->gg_g[temp[k]] = p->gg_gt[j];
      p->jptr_g[temp[k]] = i;
      temp[k]++;
    }
  free(temp);
  free(in_sigma);
  free(renum);
  return 0;
}

/* FoldedPreconditioner::BuildDiagComplex + From2x2ToCSR2x2_1/_2
   (FoldedPreconditioner.cpp:473-574,633-736) */
int orc_build_diag(orc_dfp* p, const int* ig, const int* jg, const int* idi, const int* ijg,
                   const double* di, const double* gg) {
  int n = p->n, m = p->m;
  long i, j, k;
  /* full block CSR: per row the lower blocks (ascending), the diagonal, then the upper blocks */
  long sz_jptr = (long)ig[n] * 2 + n;
  long sz_aelem = (long)ijg[ig[n]] * 2 + idi[n];
  int* iptr = (int*)malloc(sizeof(int) * ((size_t)n + 1));
  int* jptr = (int*)malloc(sizeof(int) * (size_t)sz_jptr);
  long* ijptr = (long*)malloc(sizeof(long) * ((size_t)sz_jptr + 1));
  double* aelem = (double*)malloc(sizeof(double) * (size_t)sz_aelem);
  long* col = (long*)calloc((size_t)n, sizeof(long));
  long* col2 = (long*)calloc((size_t)n, sizeof(long));
  if (!iptr || !jptr || !ijptr || !aelem || !col || !col2) return -12;
  for (i = 0; i < n; i++) {
    col[i] += ig[i + 1] - ig[i] + 1;
    col2[i] += (ijg[ig[i + 1]] - ijg[ig[i]]) + (idi[i + 1] - idi[i]);
    for (j = ig[i]; j <= ig[i + 1] - 1; j++) {
      k = jg[j];
      col[k]++;
      col2[k] += ijg[j + 1] - ijg[j];
    }
  }
  iptr[0] = 0;
  for (i = 0; i < n; i++) iptr[i + 1] = iptr[i] + (int)col[i];
  ijptr[0] = 0;
  for (i = 0; i < n; i++) ijptr[iptr[i + 1]] = ijptr[iptr[i]] + col2[i];
  for (i = 0; i < n; i++) col[i] = iptr[i];
  for (i = 0; i < n; i++) {
    for (j = ig[i]; j <= ig[i + 1] - 1; j++) {
      long c = col[i];
      int sz = ijg[j + 1] - ijg[j];
      jptr[c] = jg[j];
      ijptr[c + 1] = ijptr[c] + sz;
      for (k = 0; k < sz; k++) aelem[ijptr[c] + k] = gg[ijg[j] + k];
      col[i]++;
    }
    {
      long c = col[i];
      int sz = idi[i + 1] - idi[i];
      jptr[c] = (int)i;
      ijptr[c + 1] = ijptr[c] + sz;
      for (k = 0; k < sz; k++) aelem[ijptr[c] + k] = di[idi[i] + k];
      col[i]++;
    }
  }
  for (i = 0; i < n; i++) {
    for (j = ig[i]; j <= ig[i + 1] - 1; j++) {
      long kk = jg[j];
      long c = col[kk];
      int sz = ijg[j + 1] - ijg[j];
      jptr[c] = (int)i;
      ijptr[c + 1] = ijptr[c] + sz;
      for (long q = 0; q < sz; q++) aelem[ijptr[c] + q] = gg[ijg[j] + q];
      col[kk]++;
    }
  }
  free(col);
  free(col2);

  free(p->help);
  free(p->diag_a);
  free(p->diag_gagt);
  p->help = (double*)malloc(sizeof(double) * (2 * (size_t)m + 2));
  p->diag_a = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  p->diag_gagt = (double*)malloc(sizeof(double) * (2 * (size_t)m + 2));
  for (i = 0; i < n; i++) {
    int sz = idi[i + 1] - idi[i];
    double xr = di[idi[i]], xi = (sz == 2) ? di[idi[i] + 1] : 0.0;
    cdiv(1.0, 0.0, xr, xi, &p->diag_a[i * 2], &p->diag_a[i * 2 + 1]);
  }
  for (i = 0; i < (long)m * 2; i++) p->diag_gagt[i] = 0;
  for (i = 0; i < m; i++) {
    for (j = p->iptr_g[i]; j < p->iptr_g[i + 1]; j++) {
      k = p->jptr_g[j];
      double sr = 0.0, si = 0.0;
      for (long i2 = iptr[k]; i2 < iptr[k + 1]; i2++) {
        int j2 = jptr[i2];
        for (long i3 = p->iptr_g[i]; i3 < p->iptr_g[i + 1]; i3++) {
          if (p->jptr_g[i3] == j2) {
            long sz = ijptr[i2 + 1] - ijptr[i2];
            double xr = aelem[ijptr[i2]], xi = (sz == 2) ? aelem[ijptr[i2] + 1] : 0.0;
            /* std::complex<double> * double, then += */
            sr += xr * p->gg_g[i3];
            si += xi * p->gg_g[i3];
          }
        }
      }
      p->diag_gagt[i * 2] += sr * p->gg_g[j];
      p->diag_gagt[i * 2 + 1] += si * p->gg_g[j];
    }
  }
  for (i = 0; i < m; i++) {
    double xr = p->diag_gagt[i * 2], xi = p->diag_gagt[i * 2 + 1];
    cdiv(1.0, 0.0, xr, xi, &p->diag_gagt[i * 2], &p->diag_gagt[i * 2 + 1]);
  }
  free(iptr);
  free(jptr);
  free(ijptr);
  free(aelem);
  return 0;
}

/* FoldedPreconditioner::ApplyPreconditionerComplex (FoldedPreconditioner.cpp:578-629) */
void orc_apply_precond(const orc_dfp* p, const double* x, double* y, int T) {
  const int n = p->n, m = p->m;
  const double* da = p->diag_a;
  const double* dn = p->diag_gagt;
  double* help = p->help;
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    long lo, hi, i, j, k;
    chunk_of(n, T, t, &lo, &hi);
    for (i = lo; i < hi; i++) {
      y[i * 2] = x[i * 2] * da[i * 2] - x[i * 2 + 1] * da[i * 2 + 1];
      y[i * 2 + 1] = x[i * 2] * da[i * 2 + 1] + x[i * 2 + 1] * da[i * 2];
    }
    chunk_of((long)m * 2, T, t, &lo, &hi);
    for (i = lo; i < hi; i++) help[i] = 0;
#pragma omp barrier
    chunk_of(m, T, t, &lo, &hi);
    for (i = lo; i < hi; i++)
      for (j = p->iptr_g[i]; j < p->iptr_g[i + 1]; j++) {
        k = p->jptr_g[j];
        help[i * 2] += p->gg_g[j] * x[k * 2];
        help[i * 2 + 1] += p->gg_g[j] * x[k * 2 + 1];
      }
#pragma omp barrier
    for (i = lo; i < hi; i++) {
      double t1 = help[i * 2] * dn[i * 2] - help[i * 2 + 1] * dn[i * 2 + 1];
      double t2 = help[i * 2 + 1] * dn[i * 2] + help[i * 2] * dn[i * 2 + 1];
      help[i * 2] = t1;
      help[i * 2 + 1] = t2;
    }
#pragma omp barrier
    chunk_of(n, T, t, &lo, &hi);
    for (i = lo; i < hi; i++)
      for (j = p->iptr_gt[i]; j < p->iptr_gt[i + 1]; j++) {
        k = p->jptr_gt[j];
        y[i * 2] += p->gg_gt[j] * help[k * 2];
        y[i * 2 + 1] += p->gg_gt[j] * help[k * 2 + 1];
      }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* MRS (real variant, MRS.cpp:85-153)                                                          */

typedef struct {
  long n;
  double *y, *s, *u;
  int init;
  double residual;
} orc_mrs;

static void mrs_step(orc_mrs* M, const double* x, const double* r, int T) {
  long n = M->n, i;
  if (!M->init) {
    memcpy(M->y, x, sizeof(double) * (size_t)n);
    memcpy(M->s, r, sizeof(double) * (size_t)n);
    M->residual = norm_euclid(M->s, n, T);
    M->init = 1;
    return;
  }
  double psu[MAXT], puu[MAXT];
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    long lo, hi, q;
    double su = 0.0, uu = 0.0;
    chunk_of(n, T, t, &lo, &hi);
    for (q = lo; q < hi; q++) {
      M->u[q] = r[q] - M->s[q];
      su += M->s[q] * M->u[q];
      uu += M->u[q] * M->u[q];
    }
    psu[t] = su;
    puu[t] = uu;
  }
  double su = 0.0, uu = 0.0;
  for (int t = 0; t < T; t++) {
    su += psu[t];
    uu += puu[t];
  }
  double w;
  if (fabs(uu) < 1e-100)
    w = 0;
  else
    w = -su / uu;
  if (w < 0) {
    /* w = 0: nothing is updated and the residual keeps its previous value */
  } else if (w > 1) {
    memcpy(M->y, x, sizeof(double) * (size_t)n);
    memcpy(M->s, r, sizeof(double) * (size_t)n);
    M->residual = norm_euclid(M->s, n, T);
  } else {
    double pres[MAXT];
#pragma omp parallel num_threads(T)
    {
      int t = omp_get_thread_num();
      long lo, hi, q;
      double res = 0.0;
      chunk_of(n, T, t, &lo, &hi);
      for (q = lo; q < hi; q++) {
        M->y[q] += w * (x[q] - M->y[q]);
        M->s[q] += w * M->u[q];
        res += M->s[q] * M->s[q];
      }
      pres[t] = res;
    }
    double res = 0.0;
    for (int t = 0; t < T; t++) res += pres[t];
    M->residual = sqrt(res);
  }
  (void)i;
}

/* ------------------------------------------------------------------------------------------ */
/* PCOCR with the folded preconditioner: the request sequence of PCOCR_RCI::Run driven by     */
/* PCOCR::PCOCR_2x2_Folded, written as one loop.                                              */

static int pcocr_loop(orc_dfp* p, int n, const int* ig, const int* jg, const int* idi,
                      const int* ijg, const double* di, const double* gg, const double* pr,
                      double* x, double eps, int maxiter, int T, orc_stats* st, double* trace,
                      int trace_cap, double t_start, double* y_mrs_out) {
  const int nb = n / 2;
  const size_t bytes = sizeof(double) * (size_t)n;
  double *r = malloc(bytes), *pp = malloc(bytes), *s = malloc(bytes), *z = malloc(bytes),
         *a = malloc(bytes), *w = malloc(bytes);
  orc_mrs M;
  M.n = n;
  M.y = malloc(bytes);
  M.s = malloc(bytes);
  M.u = malloc(bytes);
  M.init = 0;
  M.residual = 0;
  double* y_omp = (T > 1) ? malloc(bytes * (size_t)(T - 1)) : NULL;
  int req = 0, iter = 0, n_spmv = 0, n_pc = 0;
  double r_old, residual_rel = 1.0;
  double beta_r = 0.0, beta_i = 0.0, zs_r, zs_i, aw_r, aw_i, al_r, al_i, zs1_r, zs1_i;
  const double eps_x0 = 1e-30, eps_zero = 1e-40;
  long i;

  /* stage 1: r = A x */
  orc_spmv(nb, ig, jg, idi, ijg, di, gg, x, r, T, y_omp);
  n_spmv++;
  /* stage 2: r = pr - r ; r_old = ||pr|| ; X0 test */
  for (i = 0; i < n; i++) r[i] = pr[i] - r[i];
  r_old = norm_euclid(pr, n, T);
  mrs_step(&M, x, r, T);
  if (norm_euclid(r, n, T) < eps_x0) {
    req = ORC_REQ_X0_NULL;
    goto done;
  }
  /* stage 3: s = M^-1 r */
  orc_apply_precond(p, r, s, T);
  n_pc++;
  /* stage 4: p = a = w = 0 ; z = A s */
  for (i = 0; i < n; i++) pp[i] = a[i] = w[i] = 0;
  orc_spmv(nb, ig, jg, idi, ijg, di, gg, s, z, T, y_omp);
  n_spmv++;
  for (;;) {
    /* stage 5 */
    iter++;
    if (iter > maxiter) {
      req = ORC_REQ_MAXITER;
      break;
    }
    cmplx_axpy(beta_r, beta_i, s, pp, pp, nb, T);
    cmplx_axpy(beta_r, beta_i, z, a, a, nb, T);
    orc_scal_cmplx(z, s, nb, T, &zs_r, &zs_i);
    orc_apply_precond(p, a, w, T);
    n_pc++;
    /* stage 6 */
    orc_scal_cmplx(a, w, nb, T, &aw_r, &aw_i);
    if (fabs(aw_r) < eps_zero && fabs(aw_i) < eps_zero) {
      req = ORC_REQ_DIV_BY_ZERO;
      break;
    }
    cdiv(zs_r, zs_i, aw_r, aw_i, &al_r, &al_i);
    cmplx_axpy(al_r, al_i, x, pp, x, nb, T);
    cmplx_axpy(-al_r, -al_i, r, a, r, nb, T);
    /* REQ_STOP_TEST: MRS then DoStopTest(double) */
    mrs_step(&M, x, r, T);
    residual_rel = M.residual / r_old;
    if (trace && iter - 1 < trace_cap) trace[iter - 1] = residual_rel;
    if (residual_rel < eps) {
      req = ORC_REQ_OK;
      break;
    }
    /* stage 7: s = s - alpha w ; z = A s */
    cmplx_axpy(-al_r, -al_i, s, w, s, nb, T);
    orc_spmv(nb, ig, jg, idi, ijg, di, gg, s, z, T, y_omp);
    n_spmv++;
    /* stage 8: beta = (z,s)_new / (z,s)_old ; the zero test looks at aw (reference behaviour) */
    zs1_r = zs_r;
    zs1_i = zs_i;
    orc_scal_cmplx(z, s, nb, T, &zs_r, &zs_i);
    if (fabs(aw_r) < eps_zero && fabs(aw_i) < eps_zero) {
      req = ORC_REQ_DIV_BY_ZERO;
      break;
    }
    cdiv(zs_r, zs_i, zs1_r, zs1_i, &beta_r, &beta_i);
  }
done:
  if (y_mrs_out) memcpy(y_mrs_out, M.y, bytes);
  if (st) {
    st->req = req;
    st->iter = iter;
    st->residual_rel = residual_rel;
    st->seconds = now_seconds() - t_start;
    st->n_spmv = n_spmv;
    st->n_precond = n_pc;
  }
  free(r);
  free(pp);
  free(s);
  free(z);
  free(a);
  free(w);
  free(M.y);
  free(M.s);
  free(M.u);
  free(y_omp);
  return req;
}

static int clamp_threads(int T) {
  if (T < 1) T = 1;
  if (T > MAXT) T = MAXT;
  return T;
}

int orc_pcocr_2x2_folded(int n, const int* ig, const int* jg, const int* idi, const int* ijg,
                         const double* di, const double* gg, const double* pr, double* x,
                         double eps, int maxiter, int kpar, int n_edges_c, int n_nodes_c,
                         const int* nvkat, const int* nver, const double* sigma3d,
                         const int* edges, const int* ig_t, const int* jg_t, const double* gg_t,
                         const int* is_node_bound, int nthreads, orc_stats* stats, double* trace,
                         int trace_cap) {
  int T = clamp_threads(nthreads);
  double t0 = now_seconds(); /* kit interval starts before the DFP setup (pcocr.cpp:61) */
  orc_dfp p;
  memset(&p, 0, sizeof(p));
  int rc = orc_build_g(&p, kpar, n_edges_c, n_nodes_c, nvkat, nver, sigma3d, edges, ig_t, jg_t,
                       gg_t, is_node_bound);
  if (rc == 0) rc = orc_build_diag(&p, ig, jg, idi, ijg, di, gg);
  if (rc != 0) {
    orc_dfp_free(&p);
    return rc;
  }
  rc = pcocr_loop(&p, n, ig, jg, idi, ijg, di, gg, pr, x, eps, maxiter, T, stats, trace,
                  trace_cap, t0, NULL);
  orc_dfp_free(&p);
  return rc;
}

int orc_pcocr_prebuilt_g(orc_dfp* p, int n, const int* ig, const int* jg, const int* idi,
                         const int* ijg, const double* di, const double* gg, const double* pr,
                         double* x, double eps, int maxiter, int nthreads, orc_stats* stats,
                         double* trace, int trace_cap, double* y_mrs_out) {
  int T = clamp_threads(nthreads);
  double t0 = now_seconds();
  int rc = orc_build_diag(p, ig, jg, idi, ijg, di, gg);
  if (rc != 0) return rc;
  return pcocr_loop(p, n, ig, jg, idi, ijg, di, gg, pr, x, eps, maxiter, T, stats, trace,
                    trace_cap, t0, y_mrs_out);
}

/* ------------------------------------------------------------------------------------------ */
/* File-driven process (HarmVect3d.cpp:118-365)                                               */

static FILE* open_ci(const char* dir, const char* const* names, const char* mode) {
  char path[4096];
  for (int k = 0; names[k]; k++) {
    snprintf(path, sizeof(path), "%s/%s", dir, names[k]);
    FILE* f = fopen(path, mode);
    if (f) return f;
  }
  return NULL;
}

static int read_bin(const char* dir, const char* const* names, void* dst, size_t elem,
                    size_t count) {
  FILE* f = open_ci(dir, names, "rb");
  if (!f) return -13;
  size_t got = fread(dst, elem, count, f);
  fclose(f);
  return got == count ? 0 : -13;
}

#define NAMES(...) ((const char* const[]){__VA_ARGS__, NULL})

int orc_cocr_fp_main(const char* dir, int nthreads_override, orc_stats* stats) {
  int rc = 0;
  FILE* f;
  /* threads: nthreads.txt capped at omp_get_max_threads()-1 (ControlOMP.cpp:41-60) */
  int nproc = 1;
  f = open_ci(dir, NAMES("nthreads.txt"), "r");
  if (f) {
    if (fscanf(f, "%d", &nproc) != 1) nproc = 1;
    fclose(f);
  }
  int maxt = omp_get_max_threads() - 1;
  if (maxt < 1) maxt = 1;
  int T = nproc <= maxt ? nproc : maxt;
  if (nthreads_override > 0) T = nthreads_override;
  T = clamp_threads(T);

  /* Read_prep_data: materials (sig3d), mu3d, dpr3d, inftry.dat, nver, nvkat, xyz, l13d */
  int n_materials = 0, tmp;
  double t1, t2;
  f = open_ci(dir, NAMES("sig3d", "Sig3d", "SIG3D"), "r");
  if (!f) return -13;
  while (fscanf(f, "%d %lf %lf", &tmp, &t1, &t2) == 3)
    if (tmp > n_materials) n_materials = tmp;
  fclose(f);
  double* sigma3d = calloc((size_t)n_materials + 1, sizeof(double));
  f = open_ci(dir, NAMES("sig3d", "Sig3d", "SIG3D"), "r");
  for (int i = 0; i < n_materials; i++) {
    int id;
    if (fscanf(f, "%d %lf %lf", &id, &t1, &t2) != 3) break;
    if (id >= 1 && id <= n_materials) sigma3d[id - 1] = t1;
  }
  fclose(f);
  f = open_ci(dir, NAMES("mu3d", "mu3D", "Mu3d", "MU3D"), "r");
  if (!f) {
    free(sigma3d);
    return -13;
  }
  fclose(f);
  f = open_ci(dir, NAMES("dpr3d", "dpr3D", "Dpr3d", "DPR3D"), "r");
  if (!f) {
    free(sigma3d);
    return -13;
  }
  fclose(f);
  int kuzlov = 0, kpar = 0, kt1 = 0;
  f = open_ci(dir, NAMES("inftry.dat"), "r");
  if (!f) {
    free(sigma3d);
    return -13;
  }
  {
    char buf[256];
    while (fscanf(f, "%255s", buf) == 1) {
      if (!strcmp(buf, "KUZLOV=")) { if (fscanf(f, "%d", &kuzlov) != 1) break; }
      else if (!strcmp(buf, "KPAR=")) { if (fscanf(f, "%d", &kpar) != 1) break; }
      else if (!strcmp(buf, "KT1=")) { if (fscanf(f, "%d", &kt1) != 1) break; }
    }
  }
  fclose(f);
  int* nver = malloc(sizeof(int) * (size_t)kpar * 14);
  int* nvkat = malloc(sizeof(int) * (size_t)kpar);
  int* l13d = malloc(sizeof(int) * ((size_t)kt1 + 1));
  rc |= read_bin(dir, NAMES("nver.dat"), nver, sizeof(int), (size_t)kpar * 14);
  rc |= read_bin(dir, NAMES("nvkat.dat"), nvkat, sizeof(int), (size_t)kpar);
  rc |= read_bin(dir, NAMES("l13d.dat", "L13d.dat", "L13D.dat"), l13d, sizeof(int), (size_t)kt1);
  for (long i = 0; i < (long)kpar * 14; i++) nver[i]--;
  for (long i = 0; i < kpar; i++) nvkat[i]--;
  for (long i = 0; i < kt1; i++) l13d[i]--;
  /* inputvec: tsize3d_.dat, nodesforedges.dat, edges.dat */
  int p_vc = 0, p_n = 0;
  f = open_ci(dir, NAMES("tsize3d_.dat"), "r");
  if (!f || fscanf(f, "%d %d", &p_vc, &p_n) != 2) rc = -13;
  if (f) fclose(f);
  int n_c = p_n - p_vc;
  int* edges = malloc(sizeof(int) * (size_t)p_n * 2 + 8);
  if (rc == 0) rc |= read_bin(dir, NAMES("nodesforedges.dat"), edges, sizeof(int), (size_t)p_n * 2);
  for (long i = 0; i < (long)p_n * 2; i++) edges[i]--;
  /* kuslau + SLAE files */
  int n = 0, maxiter = 0;
  double eps = 0;
  f = open_ci(dir, NAMES("kuslau"), "r");
  if (!f || fscanf(f, "%d %lf %d", &n, &eps, &maxiter) != 3) rc = -13;
  if (f) fclose(f);
  int *ig = NULL, *idi = NULL, *jg = NULL, *ijg = NULL, *is_node_bound = NULL, *ig_t = NULL;
  double *pr = NULL, *di = NULL, *gg = NULL, *u = NULL;
  if (rc == 0) {
    ig = malloc(sizeof(int) * ((size_t)n_c + 1));
    idi = malloc(sizeof(int) * ((size_t)n_c + 1));
    pr = malloc(sizeof(double) * (size_t)n);
    rc |= read_bin(dir, NAMES("pr"), pr, sizeof(double), (size_t)n);
    rc |= read_bin(dir, NAMES("ig"), ig, sizeof(int), (size_t)n_c + 1);
    rc |= read_bin(dir, NAMES("idi"), idi, sizeof(int), (size_t)n_c + 1);
  }
  if (rc == 0) {
    for (long i = 0; i <= n_c; i++) {
      ig[i]--;
      idi[i]--;
    }
    int nnz = ig[n_c];
    jg = malloc(sizeof(int) * ((size_t)nnz + 1));
    ijg = malloc(sizeof(int) * ((size_t)nnz + 1));
    di = malloc(sizeof(double) * ((size_t)idi[n_c] + 1));
    rc |= read_bin(dir, NAMES("di"), di, sizeof(double), (size_t)idi[n_c]);
    rc |= read_bin(dir, NAMES("jg"), jg, sizeof(int), (size_t)nnz);
    rc |= read_bin(dir, NAMES("ijg"), ijg, sizeof(int), (size_t)nnz + 1);
    if (rc == 0) {
      for (long i = 0; i < nnz; i++) jg[i]--;
      for (long i = 0; i <= nnz; i++) ijg[i]--;
      gg = malloc(sizeof(double) * ((size_t)ijg[nnz] + 1));
      rc |= read_bin(dir, NAMES("gg"), gg, sizeof(double), (size_t)ijg[nnz]);
    }
  }
  if (rc == 0) {
    u = calloc((size_t)n, sizeof(double));
    double* sig = malloc(sizeof(double) * ((size_t)n_materials + 1));
    for (int i = 0; i < n_materials; i++) sig[i] = (sigma3d[i] == 0) ? 0 : 1;
    int tsize3d = 0;
    f = open_ci(dir, NAMES("tsize3d.dat"), "r");
    if (f) {
      if (fscanf(f, "%d", &tsize3d) != 1) tsize3d = 0;
      fclose(f);
    }
    int n_nodes_c = kuzlov - tsize3d;
    ig_t = calloc((size_t)kuzlov + 1, sizeof(int));
    int* jg_t = NULL;
    double* gg_t = NULL;
    if (tsize3d) {
      rc |= read_bin(dir, NAMES("ig3d.dat"), ig_t + n_nodes_c, sizeof(int), (size_t)tsize3d + 1);
      for (int i = 0; i < n_nodes_c; i++) ig_t[i] = 1;
      for (int i = 0; i <= kuzlov; i++) ig_t[i]--;
      int nt = ig_t[kuzlov];
      jg_t = malloc(sizeof(int) * ((size_t)nt + 1));
      gg_t = malloc(sizeof(double) * ((size_t)nt + 1));
      rc |= read_bin(dir, NAMES("gg3d.dat"), gg_t, sizeof(double), (size_t)nt);
      rc |= read_bin(dir, NAMES("jg3d.dat"), jg_t, sizeof(int), (size_t)nt);
      for (int i = 0; i < nt; i++) jg_t[i]--;
    }
    is_node_bound = calloc((size_t)kuzlov + 1, sizeof(int));
    for (int i = 0; i < kt1; i++) is_node_bound[l13d[i]] = 1;
    if (rc == 0) {
      orc_stats st;
      orc_pcocr_2x2_folded(n, ig, jg, idi, ijg, di, gg, pr, u, eps, maxiter, kpar, n_c,
                           n_nodes_c, nvkat, nver, sig, edges, ig_t, jg_t, gg_t, is_node_bound,
                           T, &st, NULL, 0);
      if (stats) *stats = st;
      char path[4096];
      snprintf(path, sizeof(path), "%s/kit", dir);
      FILE* fk = fopen(path, "w");
      if (fk) {
        fprintf(fk, "r=%e\teps=%e\tkit=%d\t", st.residual_rel, eps, st.iter);
        fprintf(fk, "time=%lf sec\n", st.seconds);
        fclose(fk);
      }
      snprintf(path, sizeof(path), "%s/v3.dat", dir);
      FILE* fv = fopen(path, "wb");
      if (fv) {
        fwrite(u, sizeof(double), (size_t)n, fv);
        fclose(fv);
      } else {
        rc = -13;
      }
    }
    free(sig);
    free(jg_t);
    free(gg_t);
  }
  free(sigma3d);
  free(nver);
  free(nvkat);
  free(l13d);
  free(edges);
  free(ig);
  free(idi);
  free(jg);
  free(ijg);
  free(pr);
  free(di);
  free(gg);
  free(u);
  free(is_node_bound);
  free(ig_t);
  return rc;
}

#ifdef ORC_MAIN
int main(int argc, char** argv) {
  const char* dir = argc > 1 ? argv[1] : ".";
  int T = argc > 2 ? atoi(argv[2]) : 0;
  orc_stats st;
  memset(&st, 0, sizeof(st));
  int rc = orc_cocr_fp_main(dir, T, &st);
  if (rc != 0) {
    fprintf(stderr, "cocr_fp_oracle: input error %d\n", rc);
    return 1;
  }
  printf("req=%d iter=%d residualRel=%e time=%f\n", st.req, st.iter, st.residual_rel,
         st.seconds);
  return 0;
}
#endif
