/*
 * CPU oracle: FP64 restatement of the reference's sparse hot path.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h for who may call it, and for the
 * parity status (pinned / unpinned) of every function.
 *
 * Every function names the reference file:line it restates. Compiled with
 * -ffp-contract=off so each product is rounded before it is accumulated, as
 * in the reference's x86-64 baseline build (no FMA).
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

double orc_wtime(void) { return omp_get_wtime(); }
int orc_max_threads(void) { return omp_get_max_threads(); }

/* Planted mistakes, for the tests that show their checks catch a wrong
 * oracle: 0 = none (the default; nothing but those tests sets it). */
static int g_plant = 0;
void orc_set_plant(int which) { g_plant = which; }

/* utils/utils.h:164-186: even split of `total` items over `nthreads`,
 * the first `total % nthreads` parts one item larger. */
static void balanced_part(long total, int tid, int nthreads, long *s, long *e) {
  long wpt = total / nthreads, resid = total % nthreads;
  if (tid >= resid) {
    *s = tid * wpt + resid;
    *e = (tid + 1) * wpt + resid;
  } else {
    *s = tid * (wpt + 1);
    *e = (tid + 1) * (wpt + 1);
  }
}

/* first position p in [lo,hi) with a[p] >= v */
static long lower_bound_i(const int *a, long lo, long hi, long v) {
  while (lo < hi) {
    long mid = lo + (hi - lo) / 2;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}
/* first position p in [lo,hi) with a[p] > v */
static long upper_bound_i(const int *a, long lo, long hi, long v) {
  while (lo < hi) {
    long mid = lo + (hi - lo) / 2;
    if (a[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* ------------------------------------------------------------------ SpMV */

/* sparse_mat_op/spmv.hpp:12-29 */
void orc_spmv_serial(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *b, double *x,
                     double alpha, double beta) {
  for (int i = 0; i < n; i++) {
    double val = 0;
    const int end = ai[i + 1] - base - (g_plant == ORC_PLANT_SPMV_DROP_LAST &&
                                        ai[i + 1] > ai[i]);
    for (int j = ai[i] - base; j < end; j++)
      val += av[j] * b[aj[j] - base];
    x[i] = alpha * val + (beta == 0 ? 0 : beta * x[i]);
  }
}

/* sparse_mat_op/spmv.hpp:31-57; rows split by utils/utils.h:188-222
 * (LoadPrefixBalancedPartitionPos: lower_bound of the nnz targets in ai). */
void orc_spmv_parallel(int n, int base, const int *ai, const int *aj,
                       const double *av, const double *b, double *x,
                       double alpha, double beta, int nthreads) {
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int tid = 0; tid < nthreads; tid++) {
    long total = (long)ai[n] - ai[0];
    long wpt = total / nthreads, resid = total % nthreads;
    long t0 = tid >= resid ? tid * wpt + resid : tid * (wpt + 1);
    long t1 = tid >= resid ? (tid + 1) * wpt + resid : (tid + 1) * (wpt + 1);
    long start = tid == 0 ? 0 : lower_bound_i(ai, 0, n, t0 + ai[0]);
    long end = tid == nthreads - 1 ? n : lower_bound_i(ai, 0, n, t1 + ai[0]);
    for (long i = start; i < end; i++) {
      double val = 0;
      for (int j = ai[i] - base; j < ai[i + 1] - base; j++)
        val += av[j] * b[aj[j] - base];
      x[i] = alpha * val + (beta == 0 ? 0 : beta * x[i]);
    }
  }
}

/* sparse_mat_op/spmv.hpp:59-185. The phases separated by the reference's
 * barriers run one after another here; within a phase every "thread" works
 * on its own chunk, so the result depends only on the partition. The head
 * flags are one byte each (utils/BitVector.hpp:82-86 makes one flag per
 * char). One extra flag byte absorbs the write the reference makes at
 * index nnz for trailing empty rows (:95). */
void orc_spmv_segsum(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *b, double *x,
                     double alpha, double beta, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  const long nnz = (long)ai[n] - base;
  unsigned char *bv = (unsigned char *)calloc(nnz + 1, 1);
  double *product = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
  unsigned char *tbv = (unsigned char *)calloc(nthreads + 1, 1);
  double *tprod = (double *)calloc(nthreads + 1, sizeof(double));

  /* (1) row heads (:87-96) */
  for (int i = 0; i < n; i++) bv[ai[i] - base] = 1;

  /* (2) per-chunk product and in-chunk segmented scan (:98-116) */
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int tid = 0; tid < nthreads; tid++) {
    long start, end;
    balanced_part(nnz, tid, nthreads, &start, &end);
    for (long i = start; i < end; i++) {
      product[i] = alpha * av[i] * b[aj[i] - base];
      if (i > start && !bv[i]) {
        product[i] += product[i - 1];
        if (bv[i - 1]) bv[i] = 1;
      }
    }
    tprod[tid + 1] = 0;
    if (start < end) {
      tprod[tid + 1] = product[end - 1];
      if (bv[end - 1]) tbv[tid + 1] = 1;
    }
  }
  /* (3) serial scan of chunk tails (:118-130) */
  for (int i = 1; i <= nthreads; i++) {
    if (!tbv[i]) {
      tprod[i] += tprod[i - 1];
      if (tbv[i - 1]) tbv[i] = 1;
    }
  }
  /* (4) carry into each chunk's leading partial segment (:132-141) */
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int tid = 0; tid < nthreads; tid++) {
    long start, end;
    balanced_part(nnz, tid, nthreads, &start, &end);
    for (long i = start; i < end; i++) {
      if (!bv[i]) product[i] += tprod[tid];
      else break;
    }
  }
  /* (5) row write-back; beta is not special-cased here (:168-175) */
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int tid = 0; tid < nthreads; tid++) {
    long rs, re;
    balanced_part(n, tid, nthreads, &rs, &re);
    for (long i = rs; i < re; i++) {
      if (ai[i + 1] > ai[i]) x[i] = product[ai[i + 1] - base - 1] + beta * x[i];
      else x[i] = beta * x[i];
    }
  }
  free(bv); free(product); free(tbv); free(tprod);
}

/* sparse_mat_op/spmv.hpp:186-300 (preprocess :195-228, apply :230-293). */
void orc_spmv_albus(int n, int base, const int *ai, const int *aj,
                    const double *av, const double *b, double *x,
                    double alpha, double beta, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  const long nnz = (long)ai[n] - base;
  long *pre = (long *)malloc(sizeof(long) * (nthreads + 1));
  long *srow = (long *)malloc(sizeof(long) * (nthreads + 1));
  double *bval = (double *)calloc(2 * nthreads, sizeof(double));
  pre[0] = 0;
  long wpt = nnz / nthreads, resid = nnz % nthreads;
  for (int i = 0; i < nthreads; i++)
    pre[i + 1] = i >= resid ? ((i + 1) * wpt + resid) : ((i + 1) * (wpt + 1));
  for (int i = 0; i <= nthreads; i++)
    srow[i] = upper_bound_i(ai, 0, (long)n + 1, pre[i] + base) - 1;

#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int tid = 0; tid < nthreads; tid++) {
    long start = srow[tid], end = srow[tid + 1];
    double val;
    if (start < end) {
      val = 0;
      for (long j = pre[tid]; j < ai[start + 1] - base; j++)
        val += av[j] * b[aj[j] - base];
      start++;
      bval[2 * tid] = alpha * val;
      for (long i = start; i < end; i++) {
        val = 0;
        for (long j = ai[i] - base; j < ai[i + 1] - base; j++)
          val += av[j] * b[aj[j] - base];
        x[i] = alpha * val + (beta == 0 ? 0 : beta * x[i]);
      }
      val = 0;
      for (long j = ai[end] - base; j < pre[tid + 1]; j++)
        val += av[j] * b[aj[j] - base];
      bval[2 * tid + 1] = alpha * val;
    } else {
      val = 0;
      for (long j = pre[tid]; j < pre[tid + 1]; j++)
        val += av[j] * b[aj[j] - base];
      bval[2 * tid] = alpha * val;
      bval[2 * tid + 1] = 0;
    }
  }
  /* serial boundary fix-up (:275-292) */
  long idx = -1;
  for (int tid = 0; tid < nthreads; tid++) {
    if (srow[tid] != idx) {
      idx = srow[tid];
      x[idx] *= beta;
    }
    x[idx] += bval[2 * tid];
    if (tid == nthreads - 1) break;
    if (srow[tid + 1] != idx) {
      idx = srow[tid + 1];
      x[idx] *= beta;
    }
    x[idx] += bval[2 * tid + 1];
  }
  free(pre); free(srow); free(bval);
}

/* ----------------------------------------------------- triangular solves */

/* sparse_mat_op/triangle_solve.hpp:32-43 */
void orc_forward_sub(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *b, double *x) {
  for (int i = 0; i < n; i++) {
    double val = 0;
    for (int j = ai[i] - base; j < ai[i + 1] - base; j++)
      val += av[j] * x[aj[j] - base];
    x[i] = b[i] - val;
  }
}

/* sparse_mat_op/triangle_solve.hpp:59-71 */
void orc_backward_sub(int n, int base, const int *ai, const int *aj,
                      const double *av, const double *diag, const double *b,
                      double *x) {
  for (int i = n - 1; i >= 0; i--) {
    double val = 0;
    for (int j = ai[i] - base; j < ai[i + 1] - base; j++)
      val += av[j] * x[aj[j] - base];
    x[i] = (b[i] - val) / diag[i];
  }
}

/* sparse_mat_op/triangle_solve.hpp:87-98 */
void orc_forward_sub_t(int n, int base, const int *ai, const int *aj,
                       const double *av, const double *b, double *x) {
  memcpy(x, b, sizeof(double) * n);
  for (int i = 0; i < n; i++)
    for (int j = ai[i] - base; j < ai[i + 1] - base; j++)
      x[aj[j] - base] -= av[j] * x[i];
}

/* No counterpart in the reference: k Jacobi sweeps on (D + T) x = b with T
 * the strict triangle (either orientation) and D = diag (NULL = unit),
 * y_0 = D^-1 b, y_{s+1} = D^-1 (b - T y_s), x = y_k -- in the column-scaled
 * form the device runs: w_0 = b, w_{s+1} = b - (T D^-1) w_s with the
 * entries T_ij * (1/d_j) rounded first, x = D^-1 w_k, the last sweep scaled
 * by 1/d_i after its sum. */
void orc_trsv_jacobi(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *diag, const double *b,
                     int k, double *x) {
  double *dinv = (double *)malloc(sizeof(double) * (n ? n : 1));
  double *w = (double *)malloc(sizeof(double) * (n ? n : 1));
  double *y = (double *)malloc(sizeof(double) * (n ? n : 1));
  for (int i = 0; i < n; i++) {
    dinv[i] = diag ? 1.0 / diag[i] : 1.0;
    w[i] = b[i];
  }
  for (int s = 0; s < k; s++) {
    for (int i = 0; i < n; i++) {
      double acc = 0;
      for (int j = ai[i] - base; j < ai[i + 1] - base; j++) {
        const int c = aj[j] - base;
        const double t = diag ? av[j] * dinv[c] : av[j];
        acc += t * w[c];
      }
      y[i] = b[i] - acc;
    }
    memcpy(w, y, sizeof(double) * n);
  }
  for (int i = 0; i < n; i++) x[i] = diag ? dinv[i] * w[i] : w[i];
  free(dinv);
  free(w);
  free(y);
}

/* sparse_mat_op/matrix_utils.hpp:637-675 */
int orc_topo_sort2(int n, int base, const int *ai, const int *aj, int upper,
                   int *iperm, int *prefix) {
  int *deg = (int *)calloc(n > 0 ? n : 1, sizeof(int));
  int level = 0;
  int start = upper ? n - 1 : 0, end = upper ? -1 : n, inc = upper ? -1 : 1;
  for (int i = start; i != end; i += inc) {
    for (int j = ai[i] - base; j < ai[i + 1] - base; j++) {
      int d = deg[aj[j] - base] + 1;
      if (d > deg[i]) deg[i] = d;
    }
    if (deg[i] + 1 > level) level = deg[i] + 1;
  }
  for (int l = 0; l <= level; l++) prefix[l] = 0;
  for (int i = 0; i < n; i++) prefix[deg[i] + 1]++;
  for (int l = 1; l <= level; l++) prefix[l] += prefix[l - 1];
  for (int i = 0; i < n; i++) iperm[prefix[deg[i]]++] = i + base;
  /* rotate right by one, then prefix[0] = 0 */
  for (int l = level; l >= 1; l--) prefix[l] = prefix[l - 1];
  prefix[0] = 0;
  free(deg);
  return level;
}

/* sparse_mat_op/triangle_solve.hpp:100-145 (rows of one level are
 * independent, so the sweep order inside a level does not matter) */
void orc_level_forward(const int *iperm, const int *prefix, int lvls, int n,
                       int base, const int *ai, const int *aj,
                       const double *av, const double *b, double *x) {
  (void)n;
  for (int l = 0; l < lvls; l++) {
#pragma omp parallel for
    for (int i = prefix[l]; i < prefix[l + 1]; i++) {
      const int idx = iperm[i] - base;
      double val = 0;
      for (int j = ai[idx] - base; j < ai[idx + 1] - base; j++)
        val += av[j] * x[aj[j] - base];
      x[idx] = b[idx] - val;
    }
  }
}

void orc_level_backward(const int *iperm, const int *prefix, int lvls, int n,
                        int base, const int *ai, const int *aj,
                        const double *av, const double *diag, const double *b,
                        double *x) {
  (void)n;
  for (int l = 0; l < lvls; l++) {
#pragma omp parallel for
    for (int i = prefix[l]; i < prefix[l + 1]; i++) {
      const int idx = iperm[i] - base;
      double val = 0;
      for (int j = ai[idx] - base; j < ai[idx + 1] - base; j++)
        val += av[j] * x[aj[j] - base];
      x[idx] = (b[idx] - val) / diag[idx];
    }
  }
}

/* ------------------------------------------------------------ CSR helpers */

/* sparse_mat_op/matrix_utils.hpp:720-803 */
void orc_split_ldu(int n, int base, const int *ai, const int *aj,
                   const double *av, int *Lai, int *Laj, double *Lav,
                   double *D, int *Uai, int *Uaj, double *Uav, int *lnnz,
                   int *unnz) {
  long lp = 0, up = 0;
  Lai[0] = base;
  Uai[0] = base;
  for (int i = 0; i < n; i++) {
    long rs = ai[i] - base, re = ai[i + 1] - base;
    long mid = lower_bound_i(aj, rs, re, (long)i + base);
    int zero_diag = (mid == re || aj[mid] != i + base);
    if (D) D[i] = zero_diag ? 0 : av[mid];
    for (long j = rs; j < mid; j++) {
      if (Laj) { Laj[lp] = aj[j]; Lav[lp] = av[j]; }
      lp++;
    }
    for (long j = mid + (zero_diag ? 0 : 1); j < re; j++) {
      if (Uaj) { Uaj[up] = aj[j]; Uav[up] = av[j]; }
      up++;
    }
    Lai[i + 1] = (int)(lp + base);
    Uai[i + 1] = (int)(up + base);
  }
  if (lnnz) *lnnz = (int)lp;
  if (unnz) *unnz = (int)up;
}

/* sparse_mat_op/matrix_utils.hpp:805-879 */
void orc_split_triangle(int n, int base, const int *ai, const int *aj,
                        const double *av, int upper, int *Tai, int *Taj,
                        double *Tav, int *tnnz) {
  long p = 0;
  Tai[0] = base;
  for (int i = 0; i < n; i++) {
    long rs = ai[i] - base, re = ai[i + 1] - base;
    long s, e;
    if (upper) {
      s = lower_bound_i(aj, rs, re, (long)i + base);
      e = re;
    } else {
      s = rs;
      e = upper_bound_i(aj, rs, re, (long)i + base);
    }
    for (long j = s; j < e; j++) {
      if (Taj) { Taj[p] = aj[j]; Tav[p] = av[j]; }
      p++;
    }
    Tai[i + 1] = (int)(p + base);
  }
  if (tnnz) *tnnz = (int)p;
}

/* sparse_mat_op/matrix_utils.hpp:677-704. The reference dereferences the
 * lower_bound result even when it is the row end; here that case counts as
 * a missing diagonal. */
int orc_diagonal(int n, int base, const int *ai, const int *aj,
                 const double *av, int *diagpos, double *diag) {
  int ok = 1;
  for (int i = 0; i < n; i++) {
    long rs = ai[i] - base, re = ai[i + 1] - base;
    long mid = lower_bound_i(aj, rs, re, (long)i + base);
    if (mid == re || aj[mid] != i + base) ok = 0;
    if (diagpos) diagpos[i] = (int)(mid + base);
    if (diag && mid < re) diag[i] = av[mid];
  }
  return ok;
}

/* sparse_mat_op/matrix_utils.hpp:311-390: column counts, prefix sum, then a
 * scatter in source-row order, so every transposed row lists its columns
 * ascending. */
void orc_transpose(int rows, int cols, int base, const int *ai, const int *aj,
                   const double *av, int *ti, int *tj, double *tv) {
  long nnz = (long)ai[rows] - base;
  long *cnt = (long *)calloc((size_t)cols + 1, sizeof(long));
  for (long j = 0; j < nnz; j++) cnt[aj[j] - base + 1]++;
  for (int c = 0; c < cols; c++) cnt[c + 1] += cnt[c];
  for (int c = 0; c <= cols; c++) ti[c] = (int)(cnt[c] + base);
  for (int r = 0; r < rows; r++)
    for (long j = ai[r] - base; j < ai[r + 1] - base; j++) {
      long pos = cnt[aj[j] - base]++;
      tj[pos] = r + base;
      if (tv && av) tv[pos] = av[j];
    }
  free(cnt);
}

/* sparse_mat_op/matrix_utils.hpp:398-472 */
void orc_permute(int n, int base, const int *ai, const int *aj,
                 const double *av, const int *iperm, const int *perm,
                 int *pai, int *paj, double *pav) {
  pai[0] = base;
  for (int i = 0; i < n; i++) {
    int k = iperm ? iperm[i] - base : i;
    pai[i + 1] = pai[i] + (ai[k + 1] - ai[k]);
  }
  for (int i = 0; i < n; i++) {
    int k = iperm ? iperm[i] - base : i;
    long dst = pai[i] - base, len = ai[k + 1] - ai[k], src = ai[k] - base;
    for (long t = 0; t < len; t++) {
      int c = aj[src + t];
      paj[dst + t] = perm ? perm[c - base] : c;
      pav[dst + t] = av[src + t];
    }
    if (!perm) continue;
    /* insertion sort of the row by column (columns are distinct) */
    for (long t = 1; t < len; t++) {
      int c = paj[dst + t];
      double v = pav[dst + t];
      long u = t - 1;
      while (u >= 0 && paj[dst + u] > c) {
        paj[dst + u + 1] = paj[dst + u];
        pav[dst + u + 1] = pav[dst + u];
        u--;
      }
      paj[dst + u + 1] = c;
      pav[dst + u + 1] = v;
    }
  }
}

/* ------------------------------------------------------------------ ILU(k) */

/* mkl_wrapper/incomplete_lu.cpp:153-275. The row under construction is a
 * sorted singly linked list of (column, level), as the reference's
 * std::forward_list, held in arrays indexed by column with n as the
 * end-of-row sentinel. For level 0 the pattern is A's (:166-174); unlike
 * the reference, first_upper is filled for level 0 too (the reference
 * leaves it zero there, which its numeric phase then reads at :299). */
int orc_iluk_symbolic(int n, int base, const int *ai, const int *aj,
                      int level, int *fai, int *faj, int *first_upper) {
  int *next = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  int *lev = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  long cap = (long)ai[n] - base, pos = 0;
  int *pat = (int *)malloc(sizeof(int) * (cap > 0 ? cap : 1));
  int *plev = (int *)malloc(sizeof(int) * (cap > 0 ? cap : 1));
  long *rowptr = (long *)malloc(sizeof(long) * ((size_t)n + 1));
  rowptr[0] = 0;
  for (int i = 0; i < n; i++) {
    /* list head is the pseudo-column HEAD = n (next[n] = first column) */
    const int HEAD = n;
    int prev = HEAD;
    long list_size = ai[i + 1] - ai[i];
    for (int k = ai[i] - base; k < ai[i + 1] - base; k++) {
      int c = aj[k] - base;
      next[prev] = c;
      lev[c] = 0;
      prev = c;
    }
    next[prev] = -1; /* -1 terminates (stands for the column-n sentinel) */
    if (level > 0) {
      int cur = next[HEAD];
      for (long kk = 0; kk < list_size && cur != -1; kk++) {
        int k = cur;
        if (k >= i) break;
        int lik = lev[k];
        int e = k; /* insertion point: last list node with column <= target */
        for (long j = first_upper[k]; j < rowptr[k + 1]; j++) {
          int c = pat[j];
          int nl = lik + plev[j] + 1;
          while (next[e] != -1 && next[e] <= c) e = next[e];
          if (nl <= level) {
            if (e == c) {
              if (lev[e] > nl) lev[e] = nl;
            } else {
              next[c] = next[e];
              next[e] = c;
              lev[c] = nl;
              e = c;
              list_size++;
            }
          }
        }
        cur = next[cur];
      }
    }
    if (pos + list_size > cap) {
      while (pos + list_size > cap) cap = cap * 2 + 16;
      pat = (int *)realloc(pat, sizeof(int) * cap);
      plev = (int *)realloc(plev, sizeof(int) * cap);
    }
    for (int c = next[HEAD]; c != -1; c = next[c]) {
      pat[pos] = c;
      plev[pos] = lev[c];
      pos++;
    }
    rowptr[i + 1] = pos;
    long fu = upper_bound_i(pat, rowptr[i], rowptr[i + 1], i);
    first_upper[i] = (int)fu; /* used by later rows; 0-based position */
  }
  for (int i = 0; i <= n; i++) fai[i] = (int)(rowptr[i] + base);
  if (faj)
    for (long p = 0; p < pos; p++) faj[p] = pat[p] + base;
  free(next); free(lev); free(pat); free(plev); free(rowptr);
  return (int)pos;
}

/* mkl_wrapper/incomplete_lu.cpp:277-325 */
int orc_iluk_numeric(int n, int base, const int *ai, const int *aj,
                     const double *av, const int *fai, const int *faj,
                     const int *first_upper, double *fav) {
  double *diag = (double *)calloc(n > 0 ? n : 1, sizeof(double));
  int ok = 1;
  for (int i = 0; i < n; i++) {
    long k_idx, A_k_idx = ai[i] - base;
    long fend = fai[i + 1] - base, aend = ai[i + 1] - base;
    for (k_idx = fai[i] - base; k_idx != fend; k_idx++) {
      if (A_k_idx == aend || faj[k_idx] != aj[A_k_idx]) fav[k_idx] = 0;
      else fav[k_idx] = av[A_k_idx++];
    }
    for (k_idx = fai[i] - base; k_idx != fend && faj[k_idx] - base != i;
         k_idx++) {
      int k = faj[k_idx] - base;
      fav[k_idx] /= diag[k];
      const double aik = fav[k_idx];
      long jj = k_idx, j_idx = first_upper[k];
      long kend = fai[k + 1] - base;
      while (j_idx != kend && jj != fend) {
        if (faj[jj] == faj[j_idx]) fav[jj++] -= aik * fav[j_idx++];
        else if (faj[jj] < faj[j_idx]) jj++;
        else j_idx++;
      }
    }
    if (k_idx == fend || faj[k_idx] - base != i) ok = 0;
    else diag[i] = fav[k_idx];
  }
  free(diag);
  return ok;
}

/* ------------------------------------------------------------------ IC(k) */

/* sparse_mat_op/precond.hpp:39-143 (ICCLevelSymbolic0). llist/jk are the
 * reference's per-column linked lists of rows k whose next upper entry is
 * the column being built; NONE = INT_MAX ends a list. */
int orc_icc_symbolic(int n, int base, const int *ai, const int *aj,
                     const int *diag_pos, int lvl, int *icc_ai, int *icc_aj) {
  const int NONE = INT_MAX;
  int *llist = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  long *jk = (long *)calloc(n > 0 ? n : 1, sizeof(long));
  int *next = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  int *lev = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  long cap = (long)ai[n] - base + 16, pos = 0;
  int *pat = (int *)malloc(sizeof(int) * cap);   /* 0-based columns */
  int *plev = (int *)malloc(sizeof(int) * cap);
  long *rowptr = (long *)malloc(sizeof(long) * ((size_t)n + 1));
  for (int i = 0; i < n; i++) llist[i] = NONE;
  rowptr[0] = 0;
  for (int j = 0; j < n; j++) {
    const int HEAD = n;
    int prev = HEAD;
    long list_size = 0;
    for (long p = diag_pos[j] - base; p < ai[j + 1] - base; p++) {
      int c = aj[p] - base;
      next[prev] = c;
      lev[c] = 0;
      prev = c;
      list_size++;
    }
    next[prev] = -1;
    int k = llist[j];
    while (k < j) {
      long i_idx = jk[k]++;
      long i_end = rowptr[k + 1];
      int llist_next = llist[k];
      if (i_idx + 1 < i_end) {
        int c = pat[i_idx + 1];
        llist[k] = llist[c];
        llist[c] = k;
      }
      k = llist_next;
      int lik = plev[i_idx];
      int e = next[HEAD]; /* the merge starts at column j itself */
      for (; i_idx < i_end; i_idx++) {
        int c = pat[i_idx];
        while (next[e] != -1 && next[e] <= c) e = next[e];
        int nl = lik + plev[i_idx] + 1;
        if (nl <= lvl) {
          if (e == c) {
            if (nl < lev[e]) lev[e] = nl;
          } else {
            next[c] = next[e];
            next[e] = c;
            lev[c] = nl;
            e = c;
            list_size++;
          }
        }
      }
    }
    if (pos + list_size > cap) {
      while (pos + list_size > cap) cap = cap * 2 + 16;
      pat = (int *)realloc(pat, sizeof(int) * cap);
      plev = (int *)realloc(plev, sizeof(int) * cap);
    }
    long row_start = pos;
    for (int c = next[HEAD]; c != -1; c = next[c]) {
      pat[pos] = c;
      plev[pos] = lev[c];
      pos++;
    }
    rowptr[j + 1] = pos;
    if (pos - row_start > 1) {
      int c = pat[row_start + 1];
      llist[j] = llist[c];
      llist[c] = j;
      jk[j] = row_start + 1;
    }
  }
  for (int i = 0; i <= n; i++) icc_ai[i] = (int)(rowptr[i] + base);
  if (icc_aj)
    for (long p = 0; p < pos; p++) icc_aj[p] = pat[p] + base;
  free(llist); free(jk); free(next); free(lev); free(pat); free(plev);
  free(rowptr);
  return (int)pos;
}

/* sparse_mat_op/precond.hpp:615-693 (ICCLevelNumeric), same LIFO column
 * lists as the reference so the subtraction order is the same. */
int orc_icc_numeric(int n, int base, const int *ai, const int *aj,
                    const double *av, const int *diag_pos, double omega,
                    const int *icc_ai, const int *icc_aj, double *icc_av) {
  const int NONE = INT_MAX;
  int *llist = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  long *jk = (long *)calloc(n > 0 ? n : 1, sizeof(long));
  for (int i = 0; i < n; i++) llist[i] = NONE;
  int ok = 1;
  for (int j = 0; j < n; j++) {
    long i_idx = diag_pos[j] - base, i_end = ai[j + 1] - base;
    long ps = icc_ai[j] - base, pe = icc_ai[j + 1] - base, p = ps;
    icc_av[p++] = av[i_idx++] * (1.0 + omega);
    for (; p < pe; p++) {
      if (i_idx == i_end || icc_aj[p] != aj[i_idx]) icc_av[p] = 0;
      else icc_av[p] = av[i_idx++];
    }
    int k = llist[j];
    while (k < j) {
      long ki = jk[k]++;
      long kend = icc_ai[k + 1] - base;
      int llist_next = llist[k];
      if (ki + 1 < kend) {
        int c = icc_aj[ki + 1] - base;
        llist[k] = llist[c];
        llist[c] = k;
      }
      k = llist_next;
      const double ajk = icc_av[ki];
      long q = ps;
      while (ki < kend && q < pe) {
        if (icc_aj[ki] == icc_aj[q]) icc_av[q++] -= ajk * icc_av[ki++];
        else if (icc_aj[ki] < icc_aj[q]) ki++;
        else q++;
      }
    }
    if (pe - ps > 1) {
      int c = icc_aj[ps + 1] - base;
      llist[j] = llist[c];
      llist[c] = j;
      jk[j] = ps + 1;
    }
    if (icc_av[ps] < 0) { ok = 0; break; }
    const double aii = sqrt(icc_av[ps]);
    icc_av[ps] = aii;
    for (long q = ps + 1; q < pe; q++) icc_av[q] /= aii;
  }
  free(llist); free(jk);
  return ok;
}

/* mkl_wrapper/incomplete_cholesky.cpp:195-274, one pass of the do-while:
 * column j's bucket lists the rows k < j whose next factor entry lies in
 * column j, in the order they were appended (utils/utils.h:336-368 keeps one
 * vector per column). Visiting bucket j appends each k to the bucket of its
 * following column before k updates row j; row j joins the bucket of its
 * first off-diagonal column once its bucket is done. Additive shift on the
 * diagonal; a pivot <= 0 fails. */
typedef struct { int k; long pos; } orc_kpos;

int orc_ick_numeric(int n, int base, const int *ai, const int *aj,
                    const double *av, const int *diag_pos, double shift,
                    const int *icc_ai, const int *icc_aj, double *icc_av) {
  int nn = n > 0 ? n : 1;
  orc_kpos **bucket = (orc_kpos **)calloc(nn, sizeof(orc_kpos *));
  int *cnt = (int *)calloc(nn, sizeof(int));
  int *cap = (int *)calloc(nn, sizeof(int));
  int ok = 1;
#define ORC_PUSH(c, kk, pp)                                                   \
  do {                                                                        \
    int c_ = (c);                                                             \
    if (cnt[c_] == cap[c_]) {                                                 \
      cap[c_] = cap[c_] ? 2 * cap[c_] : 4;                                    \
      bucket[c_] = (orc_kpos *)realloc(bucket[c_], sizeof(orc_kpos) * cap[c_]); \
    }                                                                         \
    bucket[c_][cnt[c_]].k = (kk);                                             \
    bucket[c_][cnt[c_]++].pos = (pp);                                         \
  } while (0)
  for (int j = 0; j < n && ok; j++) {
    long a = diag_pos[j] - base, a_end = ai[j + 1] - base;
    long ps = icc_ai[j] - base, pe = icc_ai[j + 1] - base;
    for (long p = ps; p < pe; p++) {
      if (a == a_end || icc_aj[p] != aj[a]) icc_av[p] = 0;
      else icc_av[p] = av[a++];
    }
    icc_av[ps] += shift;
    for (int e = 0; e < cnt[j]; e++) {
      int k = bucket[j][e].k;
      long ki = bucket[j][e].pos, kend = icc_ai[k + 1] - base;
      if (ki + 1 != kend) ORC_PUSH(icc_aj[ki + 1] - base, k, ki + 1);
      const double akj = icc_av[ki];
      long q = ps;
      while (ki != kend && q != pe) {
        if (icc_aj[q] == icc_aj[ki]) icc_av[q++] -= akj * icc_av[ki++];
        else if (icc_aj[q] < icc_aj[ki]) q++;
        else ki++;
      }
    }
    free(bucket[j]);
    bucket[j] = NULL;
    cnt[j] = cap[j] = 0;
    if (pe - ps != 1) ORC_PUSH(icc_aj[ps + 1] - base, j, ps + 1);
    if (icc_av[ps] <= 0) { ok = 0; break; }
    const double d = sqrt(icc_av[ps]);
    icc_av[ps] = d;
    for (long q = ps + 1; q < pe; q++) icc_av[q] /= d;
  }
#undef ORC_PUSH
  for (int c = 0; c < nn; c++) free(bucket[c]);
  free(bucket); free(cnt); free(cap);
  return ok;
}

/* ------------------------------------------------------------- Krylov */

static double dot(int n, const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* mkl_wrapper/mkl_sparse_mat.cpp:517-533 */
void orc_rowwise_sqrt_norm(int n, int base, const int *ai, const double *av,
                           double *d) {
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = ai[i] - base; j != ai[i + 1] - base; j++) s += av[j] * av[j];
    d[i] = s > DBL_MIN ? 1. / sqrt(sqrt(s)) : 1.;
  }
}

/* mkl_wrapper/mkl_sparse_mat.cpp:535-544 */
void orc_dtad(int n, int base, const int *ai, const int *aj, double *av,
              const double *d) {
  for (int i = 0; i < n; i++)
    for (int j = ai[i] - base; j != ai[i + 1] - base; j++)
      av[j] *= d[i] * d[aj[j] - base];
}

/* ILU apply: trsv(L unit) then trsv(U non-unit) on the combined factor
 * (incomplete_lu.cpp:36-48). IC apply: trsv(U^T) then trsv(U) on the upper
 * factor incl. diagonal (incomplete_cholesky.cpp:21-29). */
void orc_prec_apply(int n, int base, const orc_prec *P, const double *r,
                    double *z) {
  if (!P || P->kind == 0) {
    memcpy(z, r, sizeof(double) * n);
    return;
  }
  const int *fi = P->fai, *fj = P->faj;
  const double *fv = P->fav;
  double *y = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  if (P->kind == 1) {
    for (int i = 0; i < n; i++) {
      double val = 0;
      for (int j = fi[i] - base; j < fi[i + 1] - base && fj[j] - base < i; j++)
        val += fv[j] * y[fj[j] - base];
      y[i] = r[i] - val;
    }
    for (int i = n - 1; i >= 0; i--) {
      double val = 0, d = 1;
      for (int j = fi[i] - base; j < fi[i + 1] - base; j++) {
        int c = fj[j] - base;
        if (c > i) val += fv[j] * z[c];
        else if (c == i) d = fv[j];
      }
      z[i] = (y[i] - val) / d;
    }
  } else {
    /* U^T y = r, column-oriented over the rows of U */
    memcpy(y, r, sizeof(double) * n);
    for (int i = 0; i < n; i++) {
      long s = fi[i] - base, e = fi[i + 1] - base;
      y[i] /= fv[s]; /* first entry of each row is the diagonal */
      for (long j = s + 1; j < e; j++) y[fj[j] - base] -= fv[j] * y[i];
    }
    for (int i = n - 1; i >= 0; i--) {
      long s = fi[i] - base, e = fi[i + 1] - base;
      double val = 0;
      for (long j = s + 1; j < e; j++) val += fv[j] * z[fj[j] - base];
      z[i] = (y[i] - val) / fv[s];
    }
  }
  free(y);
}

int orc_pcg(int n, int base, const int *ai, const int *aj, const double *av,
            const orc_prec *P, const double *b, double *x, double rel_tol,
            double abs_tol, int maxiter, int *iters, double *relres) {
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
  double *t = malloc(sizeof(double) * n);
  const double nb = sqrt(dot(n, b, b));
  int conv = 0, it = 0;
  double res = 0;
  orc_spmv_serial(n, base, ai, aj, av, x, r, 1.0, 0.0);
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
  orc_prec_apply(n, base, P, r, z);
  memcpy(p, z, sizeof(double) * n);
  double rz = dot(n, r, z);
  for (it = 1; it <= maxiter; it++) {
    orc_spmv_serial(n, base, ai, aj, av, p, q, 1.0, 0.0);
    const double alpha = rz / dot(n, p, q);
    for (int i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    /* the reference's user stopping test: an extra SpMV for the true
     * residual every iteration (mkl_solver.cpp:211-225) */
    orc_spmv_serial(n, base, ai, aj, av, x, t, 1.0, 0.0);
    for (int i = 0; i < n; i++) t[i] -= b[i];
    res = sqrt(dot(n, t, t));
    if (res / nb < rel_tol || res < abs_tol) { conv = 1; break; }
    orc_prec_apply(n, base, P, r, z);
    const double rz_new = dot(n, r, z);
    const double beta = rz_new / rz;
    rz = rz_new;
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  if (iters) *iters = conv ? it : maxiter;
  if (relres) *relres = res / nb;
  free(r); free(z); free(p); free(q); free(t);
  return conv;
}

int orc_fgmres(int n, int base, const int *ai, const int *aj,
               const double *av, const orc_prec *P, const double *b,
               double *x, int restart, double rel_tol, double abs_tol,
               int maxiter, int *iters, double *relres) {
  const int m = restart > 0 ? restart : maxiter;
  double *V = malloc(sizeof(double) * (size_t)n * (m + 1));
  double *Z = malloc(sizeof(double) * (size_t)n * m);
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = calloc(m, sizeof(double)), *sn = calloc(m, sizeof(double));
  double *g = calloc(m + 1, sizeof(double)), *y = calloc(m, sizeof(double));
  double *w = malloc(sizeof(double) * n);
  const double nb = sqrt(dot(n, b, b));
  int total = 0, conv = 0, done = 0;
  while (!done) {
    orc_spmv_serial(n, base, ai, aj, av, x, w, 1.0, 0.0);
    for (int i = 0; i < n; i++) w[i] = b[i] - w[i];
    double beta = sqrt(dot(n, w, w));
    if (beta == 0) { conv = 1; break; }
    for (int i = 0; i < n; i++) V[i] = w[i] / beta;
    memset(g, 0, sizeof(double) * (m + 1));
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m; j++) {
      double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n;
      orc_prec_apply(n, base, P, vj, zj);
      orc_spmv_serial(n, base, ai, aj, av, zj, w, 1.0, 0.0);
      for (int i = 0; i <= j; i++) {
        const double *vi = V + (size_t)i * n;
        double h = dot(n, w, vi);
        H[(size_t)i * m + j] = h;
        for (int t = 0; t < n; t++) w[t] -= h * vi[t];
      }
      double hn = sqrt(dot(n, w, w));
      H[(size_t)(j + 1) * m + j] = hn;
      for (int i = 0; i < j; i++) {
        double a = H[(size_t)i * m + j], c = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a + sn[i] * c;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
      }
      double a = H[(size_t)j * m + j], c = H[(size_t)(j + 1) * m + j];
      double rr = sqrt(a * a + c * c);
      cs[j] = rr == 0 ? 1 : a / rr;
      sn[j] = rr == 0 ? 0 : c / rr;
      H[(size_t)j * m + j] = rr;
      H[(size_t)(j + 1) * m + j] = 0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      total++;
      k = j + 1;
      const double est = fabs(g[j + 1]);
      if (est / nb < rel_tol || est < abs_tol) { conv = 1; done = 1; break; }
      if (hn < 1.0e-16) { conv = 1; done = 1; break; }
      if (total >= maxiter) { done = 1; break; }
      double *vn = V + (size_t)(j + 1) * n;
      for (int t = 0; t < n; t++) vn[t] = w[t] / hn;
    }
    for (int i = k - 1; i >= 0; i--) {
      double s = g[i];
      for (int l = i + 1; l < k; l++) s -= H[(size_t)i * m + l] * y[l];
      y[i] = s / H[(size_t)i * m + i];
    }
    for (int l = 0; l < k; l++) {
      const double *zl = Z + (size_t)l * n;
      for (int t = 0; t < n; t++) x[t] += y[l] * zl[t];
    }
  }
  orc_spmv_serial(n, base, ai, aj, av, x, w, 1.0, 0.0);
  for (int i = 0; i < n; i++) w[i] = b[i] - w[i];
  if (iters) *iters = total;
  if (relres) *relres = nb > 0 ? sqrt(dot(n, w, w)) / nb : 0;
  free(V); free(Z); free(H); free(cs); free(sn); free(g); free(y); free(w);
  return conv;
}

/* ------------------------------------------------------------------ RCM */

/* reordering/BFS.cpp:10-57 (BFS_Fn): FIFO BFS, levels[] = distance, -1 if
 * unreached; returns the number of levels (height). */
int orc_bfs_levels(int n, int base, const int *ai, const int *aj, int source,
                   int *levels) {
  int *queue = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int head = 0, tail = 0, height = 0;
  for (int i = 0; i < n; i++) levels[i] = -1;
  queue[tail++] = source - base;
  levels[source - base] = 0;
  while (head < tail) {
    int u = queue[head++];
    for (int k = ai[u] - base; k < ai[u + 1] - base; k++) {
      int v = aj[k] - base;
      if (levels[v] == -1) {
        if (height < levels[u] + 1) height = levels[u] + 1;
        levels[v] = height;
        queue[tail++] = v;
      }
    }
  }
  free(queue);
  return height + 1;
}

/* reordering/BFS.cpp:59-180 (PBFS_Fn<true,false>) as PseudoDiameter uses
 * it: frontier sizes only, no level array. *width is the BFS object's
 * member: max'ed with every frontier size (the final empty one included)
 * and never reset between calls on the same object. Returns 0 when a
 * frontier reaches `shortcut`. The last level comes back as a set (its
 * order never matters to the caller). */
static int pbfs_lastlevel(int n, int base, const int *ai, const int *aj,
                          int source, int shortcut, int *height, int *width,
                          int *last, int *nlast, unsigned char *visited,
                          int *bufa, int *bufb) {
  memset(visited, 0, (size_t)n);
  int *front = bufa, *nxt = bufb, *prev = NULL;
  int nf = 1, nprev = 0, h = 0;
  front[0] = source - base;
  visited[source - base] = 1;
  while (1) {
    if (nf > *width) *width = nf;
    h++;
    if (nf == 0) {
      for (int q = 0; q < nprev; q++) last[q] = prev[q] + base;
      *nlast = nprev;
      break;
    }
    if (nf >= shortcut) {
      *height = h - 1;
      return 0;
    }
    int nn = 0;
    for (int q = 0; q < nf; q++) {
      int u = front[q];
      for (int k = ai[u] - base; k < ai[u + 1] - base; k++) {
        int v = aj[k] - base;
        if (!visited[v]) {
          visited[v] = 1;
          nxt[nn++] = v;
        }
      }
    }
    prev = front;
    nprev = nf;
    front = nxt;
    nxt = prev;
    nf = nn;
  }
  *height = h - 1;
  return 1;
}

/* reordering/Reordering.h:73-136 (PseudoDiameter); mutates degrees. */
static void pseudo_diameter(int n, int base, const int *ai, const int *aj,
                            int *degrees, const int *view, int nview,
                            int *source_out, int *target_out) {
  int source = -1, best = INT_MAX;
  for (int q = 0; q < nview; q++)
    if (degrees[view[q] - base] < best) {
      best = degrees[view[q] - base];
      source = view[q];
    }
  int target = -1;
  unsigned char *visited = (unsigned char *)malloc((size_t)n);
  int *bufa = (int *)malloc(sizeof(int) * n), *bufb = (int *)malloc(sizeof(int) * n);
  int *last = (int *)malloc(sizeof(int) * n), *last2 = (int *)malloc(sizeof(int) * n);
  int chosen[5];
  int forwardWidth = 0, backwardWidth = INT_MAX, diameter = 0;
  while (target == -1) {
    int nchosen = 0, width = 0, height = 0, nlast = 0;
    pbfs_lastlevel(n, base, ai, aj, source, INT_MAX, &height, &width, last,
                   &nlast, visited, bufa, bufb);
    diameter = height;
    forwardWidth = width;
    /* "first five" strategy */
    while (nchosen < 5) {
      int minDeg = INT_MAX, sel = -1;
      for (int q = 0; q < nlast; q++) {
        int i = last[q];
        if (degrees[i - base] < minDeg) {
          minDeg = degrees[i - base];
          sel = i;
        } else if (degrees[i - base] == minDeg) {
          if (i < sel) sel = i;
        }
      }
      if (minDeg == INT_MAX) break;
      chosen[nchosen++] = sel;
      degrees[sel - base] = INT_MAX;
      for (int k = ai[sel - base] - base; k < ai[sel - base + 1] - base; k++)
        degrees[aj[k] - base] = INT_MAX;
    }
    backwardWidth = INT_MAX;
    for (int c = 0; c < nchosen; c++) {
      int i = chosen[c], h2 = 0, nl2 = 0;
      if (g_plant == ORC_PLANT_RCM_WIDTH_RESET) width = 0;
      if (!pbfs_lastlevel(n, base, ai, aj, i, backwardWidth, &h2, &width,
                          last2, &nl2, visited, bufa, bufb))
        continue;
      if (diameter < h2 && width < backwardWidth) {
        source = i;
        break;
      } else if (width < backwardWidth) {
        backwardWidth = width;
        target = i;
      }
    }
    if (nchosen == 0) break; /* the reference would loop forever here */
  }
  if (forwardWidth > backwardWidth) {
    int t = source;
    source = target;
    target = t;
  }
  *source_out = source;
  *target_out = target;
  free(visited); free(bufa); free(bufb); free(last); free(last2);
}

/* reordering/UnionFind.cpp:82-129 (ParUnionFindRem) and :217-293
 * (ComponentsStat): parents always point to a larger index, so each
 * component's root is its largest node; components come in ascending root
 * order, members ascending. */
static int rem_find(int *par, int x) {
  while (x != par[x]) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

void orc_rcm(int n, int base, const int *ai, const int *aj, int *iperm,
             int *perm) {
  int *deg = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *deg0 = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *par = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    deg[i] = deg0[i] = ai[i + 1] - ai[i];
    par[i] = i;
  }
  for (int i = 0; i < n; i++)
    for (int j = ai[i] - base; j < ai[i + 1] - base; j++) {
      int x = i, y = aj[j] - base;
      while (1) {
        int px = par[x], py = par[y];
        if (px == py) break;
        if (py < px) { int t = x; x = y; y = t; t = px; px = py; py = t; }
        if (x == px) { par[x] = py; break; }
        par[x] = py;
        x = px;
      }
    }
  int *root = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *cidx = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int ncomp = 0;
  for (int i = 0; i < n; i++) root[i] = rem_find(par, i);
  for (int i = 0; i < n; i++) cidx[i] = -1;
  int *compOf = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++)
    if (root[i] == i) cidx[i] = ncomp++;
  int *cpre = (int *)calloc((size_t)ncomp + 1, sizeof(int));
  for (int i = 0; i < n; i++) cpre[cidx[root[i]] + 1]++;
  for (int c = 0; c < ncomp; c++) cpre[c + 1] += cpre[c];
  int *fill = (int *)malloc(sizeof(int) * ((size_t)ncomp + 1));
  memcpy(fill, cpre, sizeof(int) * ((size_t)ncomp + 1));
  int *sorted = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) sorted[fill[cidx[root[i]]]++] = i + base;
  (void)compOf;

  int *levels = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *prefix = (int *)malloc(sizeof(int) * ((size_t)n + 2));
  int *children = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int c = 0; c < ncomp; c++) {
    int off = cpre[c], sz = cpre[c + 1] - cpre[c];
    if (sz == 1) { iperm[off] = sorted[off]; continue; }
    if (sz == 2) {
      iperm[off] = sorted[off];
      iperm[off + 1] = sorted[off + 1];
      continue;
    }
    int source, target;
    pseudo_diameter(n, base, ai, aj, deg, sorted + off, sz, &source, &target);
    int e = off;
    iperm[e++] = source;
    int height = orc_bfs_levels(n, base, ai, aj, source, levels);
    for (int l = 0; l <= height; l++) prefix[l] = 0;
    for (int p = off; p < off + sz; p++) prefix[levels[sorted[p] - base] + 1]++;
    for (int l = 0; l < height; l++) prefix[l + 1] += prefix[l];
    for (int l = 0; l < height; l++) {
      for (int r = prefix[l]; r != prefix[l + 1]; r++) {
        int nc = 0;
        int u = iperm[r + off] - base;
        for (int j = ai[u] - base; j != ai[u + 1] - base; j++) {
          int v = aj[j] - base;
          if (levels[v] == l + 1) {
            children[nc++] = v;
            levels[v] = -1;
          }
        }
        /* insertion sort by (degree, index) (Reordering.cpp:117-122),
         * with the degrees PseudoDiameter overwrote */
        const int *sd = g_plant == ORC_PLANT_RCM_ORIGINAL_DEGREES ? deg0 : deg;
        for (int a = 1; a < nc; a++) {
          int v = children[a], t = a - 1;
          while (t >= 0 && (sd[children[t]] > sd[v] ||
                            (sd[children[t]] == sd[v] && children[t] > v))) {
            children[t + 1] = children[t];
            t--;
          }
          children[t + 1] = v;
        }
        for (int a = 0; a < nc; a++) iperm[e + a] = children[a] + base;
        e += nc;
      }
    }
  }
  for (int i = 0, j = n - 1; i < j; i++, j--) {
    int t = iperm[i]; iperm[i] = iperm[j]; iperm[j] = t;
  }
  for (int i = 0; i < n; i++) perm[iperm[i] - base] = i + base;
  free(deg); free(deg0); free(par); free(root); free(cidx); free(cpre); free(fill);
  free(sorted); free(levels); free(prefix); free(children); free(compOf);
}
