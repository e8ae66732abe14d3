/*
 * oracle_spai.c -- CPU restatement of the learned sparse approximate inverse
 * (no reference counterpart: the reference's sparse approximate inverses,
 * neumann_preconditioner.hpp and least_squares_preconditioner.hpp, are
 * polynomial, not learned).  TEST INFRASTRUCTURE ONLY (see
 * sparkit_oracle.h).
 *
 * Row slots.  Row i must hold its diagonal, a_ii > 0, and at most 4 entries
 * on each side of it.  Slot 4 is the diagonal; the entries left of it, from
 * the nearest outwards, take slots 3, 2, 1, 0; those right of it take slots
 * 5, 6, 7, 8 (for an interior row of a 9-point stencil in natural order the
 * slots are the stencil positions SW S SE W C E NW N NE).
 *
 * Features (24 per row, computed in double, stored as float), for the eight
 * off-diagonal slots k = 0..7 (slots 0 1 2 3 5 6 7 8), column j, zero when
 * the slot is empty:
 *   x[k]      = a_ij / a_ii
 *   x[8 + k]  = (a_jj - a_ii) / (a_jj + a_ii)
 *   x[16 + k] = rowsum_j / a_jj,  rowsum_j = sum of row j in storage order
 *
 * Network: y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3 with 24 -> 64 ->
 * 64 -> 9; the GEMM operands (x, weights, hidden activations, backward
 * deltas) are rounded to bfloat16 and accumulated in fp32.
 *
 * M_i,slot = (delta_{slot,4} + y_slot) / a_ii on row i's slots (zero
 * weights give the Jacobi preconditioner).
 *
 * Loss = (1/n) ||I - M A||_F^2, row by row: (M A)_i = sum_slots m_s A_{j_s,:}
 * over the union of the neighbour rows' columns (a merge of sorted rows).
 *
 * Parameter layout (canonical, 6345 values): W1[24][64] b1[64] W2[64][64]
 * b2[64] W3[64][9] b3[9], weights indexed [in][out].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sparkit_oracle.h"

#define SPAI_F 24
#define SPAI_H 64
#define SPAI_O 9
#define SPAI_W1 0
#define SPAI_B1 (SPAI_W1 + SPAI_F * SPAI_H)
#define SPAI_W2 (SPAI_B1 + SPAI_H)
#define SPAI_B2 (SPAI_W2 + SPAI_H * SPAI_H)
#define SPAI_W3 (SPAI_B2 + SPAI_H)
#define SPAI_B3 (SPAI_W3 + SPAI_H * SPAI_O)
#define SPAI_NPARAM (SPAI_B3 + SPAI_O)

/* diagonal positions and row sums; validates the slot rules */
static int spai_prepare(const orc_csr_d* A, double** diag_out,
                        double** rsum_out) {
  const int64_t n = A->nrows;
  const int64_t *rp = A->row_ptr, *ci = A->col_ind;
  const double* v = A->vals;
  double* diag = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double* rsum = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int st = ORC_OK;
  for (int64_t i = 0; i < n && st == ORC_OK; ++i) {
    int64_t dpos = -1;
    double s = 0.0;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
      if (ci[p] == i) dpos = p;
      s += v[p];
    }
    if (dpos < 0 || !(v[dpos] > 0.0) || dpos - rp[i] > 4 ||
        rp[i + 1] - dpos - 1 > 4)
      st = ORC_INVALID_ARGUMENT;
    else
      diag[i] = v[dpos];
    rsum[i] = s;
  }
  if (st != ORC_OK || A->nrows != A->ncols) {
    free(diag);
    free(rsum);
    return ORC_INVALID_ARGUMENT;
  }
  *diag_out = diag;
  *rsum_out = rsum;
  return ORC_OK;
}

static void spai_row(const orc_csr_d* A, const double* diag, const double* rsum,
                     int64_t i, double* x, int64_t* pos) {
  const int64_t *rp = A->row_ptr, *ci = A->col_ind;
  const double* v = A->vals;
  int64_t dpos = rp[i];
  while (ci[dpos] != i) ++dpos;
  for (int s = 0; s < SPAI_O; ++s) pos[s] = -1;
  pos[4] = dpos;
  for (int64_t k = 0; dpos - 1 - k >= rp[i]; ++k) pos[3 - k] = dpos - 1 - k;
  for (int64_t k = 0; dpos + 1 + k < rp[i + 1]; ++k) pos[5 + k] = dpos + 1 + k;
  static const int off[8] = {0, 1, 2, 3, 5, 6, 7, 8};
  const double di = diag[i];
  for (int k = 0; k < 8; ++k) {
    int64_t p = pos[off[k]];
    if (p < 0) {
      x[k] = x[8 + k] = x[16 + k] = 0.0;
      continue;
    }
    const int64_t j = ci[p];
    x[k] = v[p] / di;
    x[8 + k] = (diag[j] - di) / (diag[j] + di);
    x[16 + k] = rsum[j] / diag[j];
  }
}

/* ||e_i^T - m^T A_{slots,:}||^2 by a merge of the slot rows; g_s receives
 * d/dm_s = -2 sum_c res_c a_{j_s,c} */
static double spai_row_residual(const orc_csr_d* A, int64_t i,
                                const int64_t* pos, const double* m,
                                double* g) {
  const int64_t *rp = A->row_ptr, *ci = A->col_ind;
  const double* v = A->vals;
  int64_t p[SPAI_O], e[SPAI_O];
  for (int s = 0; s < SPAI_O; ++s) {
    if (pos[s] < 0) {
      p[s] = e[s] = 0;
      continue;
    }
    int64_t j = ci[pos[s]];
    p[s] = rp[j];
    e[s] = rp[j + 1];
  }
  double loss = 0.0;
  for (;;) {
    int64_t c = INT64_MAX;
    for (int s = 0; s < SPAI_O; ++s)
      if (p[s] < e[s] && ci[p[s]] < c) c = ci[p[s]];
    if (c == INT64_MAX) break;
    double t = 0.0;
    for (int s = 0; s < SPAI_O; ++s)
      if (p[s] < e[s] && ci[p[s]] == c) t += m[s] * v[p[s]];
    double res = (c == i ? 1.0 : 0.0) - t;
    loss += res * res;
    for (int s = 0; s < SPAI_O; ++s)
      if (p[s] < e[s] && ci[p[s]] == c) {
        g[s] += -2.0 * res * v[p[s]];
        ++p[s];
      }
  }
  return loss;
}

/* position of entry (j, i) for entry p = (i, j), by a linear scan of row j;
 * -1 when the pattern has no such entry */
static int64_t spai_mirror(const orc_csr_d* A, int64_t i, int64_t p) {
  const int64_t j = A->col_ind[p];
  for (int64_t q = A->row_ptr[j]; q < A->row_ptr[j + 1]; ++q)
    if (A->col_ind[q] == i) return q;
  return -1;
}

int orc_spai_features(const orc_csr_d* A, float* X) {
  double *diag, *rsum;
  int st = spai_prepare(A, &diag, &rsum);
  if (st != ORC_OK) return st;
  for (int64_t i = 0; i < A->nrows; ++i) {
    double x[SPAI_F];
    int64_t pos[SPAI_O];
    spai_row(A, diag, rsum, i, x, pos);
    for (int f = 0; f < SPAI_F; ++f) X[i * SPAI_F + f] = (float)x[f];
  }
  free(diag);
  free(rsum);
  return ORC_OK;
}

float orc_bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return x; /* NaN */
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  memcpy(&x, &u, 4);
  return x;
}

#define REAL float
#define RND(v) orc_bf16_round(v)
#define SPAI_FN(name) name##_bf16
#include "oracle_spai.inc"
#undef REAL
#undef RND
#undef SPAI_FN

#define REAL double
#define RND(v) (v)
#define SPAI_FN(name) name##_exact
#include "oracle_spai.inc"
#undef REAL
#undef RND
#undef SPAI_FN
