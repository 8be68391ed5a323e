/*
 * CPU restatement of the matvec over the flattened traversal -- test
 * infrastructure and the CPU baseline (bench.py's cpu_baseline leg); the
 * product never links it.
 *
 * The per-leaf step of the reference (FEM/include/matvec.h:380-522) applied to
 * the maps that dkt_da_traversal and dkt_da_stale_terms export (the traversal
 * of matvec_rec, matvec.h:243-548, flattened once per mesh):
 *   - gather the present ranks (slot >= 0) of the leaf;
 *   - a hanging leaf interpolates its missing ranks from its parent-level
 *     nodes (pslot; absent parent ranks read zero) with the child's
 *     tensor-product interpolation (RefElement::IKD_Parent2Child,
 *     FEM/include/refel.h:214-240), plus the stale parent-buffer terms;
 *   - v = 2^(-level * hp) A_ref u with the closed-form Q_p operator (op 0:
 *     stiffness, hp = dim - 2; op 1: mass, hp = dim; op 2: identity);
 *   - present ranks add v to their nodes; a hanging leaf sends its missing
 *     ranks' outputs through the transposed interpolation
 *     (IKD_Child2Parent, refel.h:243-271) to the parent-level nodes at ranks
 *     the leaf lacks (matvec.h:482-522).
 * Threads: leaves are cut into nthreads contiguous ranges, each accumulated
 * into its own output vector, summed in thread order (deterministic for a
 * given thread count).  Matches the reference's outputs to rounding (the
 * summation order differs), which tests/test_oracle.py checks on every golden.
 */
#include "plan_cpu.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MAXN 81 /* (p + 1)^dim for p <= 2, dim <= 4 */

/* Lagrange basis k of nodes {0, 1/p, ..., 1} at t. */
static double lagrange(int p, int k, double t) {
  double v = 1.0;
  for (int j = 0; j <= p; j++)
    if (j != k) v *= (t - (double)j / p) / ((double)(k - j) / p);
  return v;
}

static void one_d(int p, double *K, double *M) {
  if (p == 1) {
    const double k[4] = {1, -1, -1, 1}, m[4] = {1.0 / 3, 1.0 / 6, 1.0 / 6, 1.0 / 3};
    memcpy(K, k, sizeof k);
    memcpy(M, m, sizeof m);
  } else {
    const double k[9] = {7.0 / 3, -8.0 / 3, 1.0 / 3, -8.0 / 3, 16.0 / 3, -8.0 / 3, 1.0 / 3, -8.0 / 3, 7.0 / 3};
    const double m[9] = {4.0 / 30, 2.0 / 30, -1.0 / 30, 2.0 / 30, 16.0 / 30, 2.0 / 30, -1.0 / 30, 2.0 / 30, 4.0 / 30};
    memcpy(K, k, sizeof k);
    memcpy(M, m, sizeof m);
  }
}

/* A_ref[i][j], lexicographic ranks with axis 0 fastest. */
static void ref_matrix(int dim, int p, int op, double *A, int npe) {
  double K[9], M[9];
  one_d(p, K, M);
  const int n = p + 1;
  for (int i = 0; i < npe; i++)
    for (int j = 0; j < npe; j++) {
      double s = 0.0;
      if (op == 2) {
        s = i == j ? 1.0 : 0.0;
      } else {
        const int nterms = op == 0 ? dim : 1;
        for (int t = 0; t < nterms; t++) {
          double prod = 1.0;
          int a = i, b = j;
          for (int d = 0; d < dim; d++, a /= n, b /= n) {
            const double *m = (op == 0 && d == t) ? K : M;
            prod *= m[(a % n) * n + (b % n)];
          }
          s += prod;
        }
      }
      A[i * npe + j] = s;
    }
}

/* W_c[r][s]: weight of parent rank s in child rank r for child c. */
static void interp_matrix(int dim, int p, int c, double *W, int npe) {
  const int n = p + 1;
  for (int r = 0; r < npe; r++)
    for (int s = 0; s < npe; s++) {
      double w = 1.0;
      int a = r, b = s;
      for (int d = 0; d < dim; d++, a /= n, b /= n) {
        const double t = 0.5 * (((c >> d) & 1) + (double)(a % n) / p);
        w *= lagrange(p, b % n, t);
      }
      W[r * npe + s] = w;
    }
}

int oracle_plan_matvec(int dim, int p, size_t n_nodes, size_t L, const uint32_t *leaf_cells, const uint8_t *child,
                       const int32_t *slot, const int32_t *hang_of, const int32_t *pslot, const int64_t *st_off,
                       const uint8_t *st_rank, const int32_t *st_node, const double *st_w, int op, const double *x,
                       double *y, int nthreads) {
  if (dim < 2 || dim > 4 || p < 1 || p > 2 || op < 0 || op > 2 || nthreads < 1) return 1;
  int npe = 1;
  for (int d = 0; d < dim; d++) npe *= p + 1;
  if (npe > MAXN) return 1;
  const int hp = op == 0 ? dim - 2 : (op == 1 ? dim : 0);
  double *A = malloc(sizeof(double) * npe * npe);
  double *W = malloc(sizeof(double) * (1 << dim) * npe * npe);
  double *part = calloc((size_t)nthreads * n_nodes, sizeof(double));
  if (!A || !W || !part) {
    free(A);
    free(W);
    free(part);
    return 2;
  }
  ref_matrix(dim, p, op, A, npe);
  for (int c = 0; c < (1 << dim); c++) interp_matrix(dim, p, c, W + (size_t)c * npe * npe, npe);
#pragma omp parallel num_threads(nthreads)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const size_t e0 = L * t / nt, e1 = L * (t + 1) / nt;
    double *yt = part + (size_t)t * n_nodes;
    double u[MAXN], v[MAXN], xp[MAXN], w[MAXN];
    for (size_t e = e0; e < e1; e++) {
      const int32_t *s = slot + e * npe;
      const int h = hang_of[e];
      const double *Wc = W + (size_t)child[e] * npe * npe;
      for (int r = 0; r < npe; r++) u[r] = s[r] >= 0 ? x[s[r]] : 0.0;
      if (h >= 0) {
        const int32_t *ps = pslot + (size_t)h * npe;
        for (int k = 0; k < npe; k++) xp[k] = ps[k] >= 0 ? x[ps[k]] : 0.0;
        for (int r = 0; r < npe; r++) {
          if (s[r] >= 0) continue;
          double acc = 0.0;
          for (int k = 0; k < npe; k++) acc += Wc[r * npe + k] * xp[k];
          u[r] = acc;
        }
        for (int64_t k = st_off[h]; k < st_off[h + 1]; k++) u[st_rank[k]] += st_w[k] * x[st_node[k]];
      }
      const int lev = (int)leaf_cells[e * (dim + 1) + dim];
      const double scale = ldexp(1.0, -lev * hp);
      for (int i = 0; i < npe; i++) {
        double acc = 0.0;
        for (int j = 0; j < npe; j++) acc += A[i * npe + j] * u[j];
        v[i] = scale * acc;
      }
      for (int r = 0; r < npe; r++)
        if (s[r] >= 0) yt[s[r]] += v[r];
      if (h >= 0) {
        const int32_t *ps = pslot + (size_t)h * npe;
        for (int k = 0; k < npe; k++) {
          if (ps[k] < 0 || s[k] >= 0) continue;
          double acc = 0.0;
          for (int r = 0; r < npe; r++)
            if (s[r] < 0) acc += Wc[r * npe + k] * v[r];
          w[k] = acc;
          yt[ps[k]] += w[k];
        }
      }
    }
#pragma omp barrier
#pragma omp for schedule(static)
    for (size_t i = 0; i < n_nodes; i++) {
      double acc = 0.0;
      for (int k = 0; k < nt; k++) acc += part[(size_t)k * n_nodes + i];
      y[i] = acc;
    }
  }
  free(A);
  free(W);
  free(part);
  return 0;
}
