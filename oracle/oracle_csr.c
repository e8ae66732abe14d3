/*
 * oracle_csr.c -- CSR construction, SpMV, Jacobi, ILU(0), triangular solves
 * and the reference generators, restated in C.  TEST INFRASTRUCTURE ONLY
 * (see sparkit_oracle.h).
 */
#include "sparkit_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Sorting columns inside one row (insertion sort for the short rows of
 * stencils, qsort above a threshold).  Only the sorted, de-duplicated result
 * matters for bit-exactness, not the sorting algorithm. */
static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

static void sort_row(int64_t* c, int64_t len) {
  if (len > 32) {
    qsort(c, (size_t)len, sizeof(int64_t), cmp_i64);
    return;
  }
  for (int64_t i = 1; i < len; ++i) {
    int64_t v = c[i];
    int64_t j = i - 1;
    while (j >= 0 && c[j] > v) {
      c[j + 1] = c[j];
      --j;
    }
    c[j + 1] = v;
  }
}

/* Compressed_row_sparsity_Impl.cpp:10-40.  The reference sorts the (row,
 * col) list and calls std::unique; here a counting sort by row is followed
 * by a sort of each row's columns and an in-place unique, which yields the
 * same sorted unique sequence.  Index.cpp:8 rejects negative indices,
 * Shape.cpp:8-10 rejects shapes with rows <= 1 or cols <= 1. */
int orc_csr_pattern(int64_t nrows, int64_t ncols, int64_t nidx,
                    const int64_t* rows, const int64_t* cols,
                    int64_t* row_ptr, int64_t* col_ind, int64_t* nnz_out) {
  if (nrows <= 1 || ncols <= 1) return ORC_LOGIC_ERROR;
  for (int64_t k = 0; k < nidx; ++k) {
    if (rows[k] < 0 || cols[k] < 0) return ORC_LOGIC_ERROR;
    /* the reference indexes row_ptr_ by row without a bound check (UB);
     * out-of-shape indices are rejected here instead */
    if (rows[k] >= nrows || cols[k] >= ncols) return ORC_INVALID_ARGUMENT;
  }
  int64_t* cnt = (int64_t*)calloc((size_t)nrows + 1, sizeof(int64_t));
  if (!cnt) return ORC_NO_MEMORY;
  for (int64_t k = 0; k < nidx; ++k) cnt[rows[k] + 1]++;
  for (int64_t r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
  int64_t* fill = (int64_t*)malloc(((size_t)nrows + 1) * sizeof(int64_t));
  if (!fill) {
    free(cnt);
    return ORC_NO_MEMORY;
  }
  memcpy(fill, cnt, ((size_t)nrows + 1) * sizeof(int64_t));
  /* scatter columns by row into col_ind (capacity nidx) */
  for (int64_t k = 0; k < nidx; ++k) col_ind[fill[rows[k]]++] = cols[k];
  /* sort + unique per row, compacting in place */
  int64_t out = 0;
  row_ptr[0] = 0;
  for (int64_t r = 0; r < nrows; ++r) {
    int64_t b = cnt[r], e = cnt[r + 1];
    sort_row(col_ind + b, e - b);
    for (int64_t p = b; p < e; ++p) {
      if (p > b && col_ind[p] == col_ind[p - 1]) continue;
      col_ind[out++] = col_ind[p];
    }
    row_ptr[r + 1] = out;
  }
  *nnz_out = out;
  free(fill);
  free(cnt);
  return ORC_OK;
}

static int64_t find_in_row(const int64_t* row_ptr, const int64_t* col_ind,
                           int64_t r, int64_t c) {
  int64_t lo = row_ptr[r], hi = row_ptr[r + 1];
  while (lo < hi) { /* std::lower_bound */
    int64_t mid = lo + (hi - lo) / 2;
    if (col_ind[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < row_ptr[r + 1] && col_ind[lo] == c) return lo;
  return -1;
}

/* matgen.hpp:25-52: values start at 0 and each entry, in input order, is
 * written at its position (a later duplicate overwrites an earlier one). */
int orc_make_matrix(int64_t nrows, int64_t ncols, int64_t nent,
                    const int64_t* rows, const int64_t* cols, const double* v,
                    int64_t* row_ptr, int64_t* col_ind, double* vals,
                    int64_t* nnz_out) {
  int st = orc_csr_pattern(nrows, ncols, nent, rows, cols, row_ptr, col_ind,
                           nnz_out);
  if (st != ORC_OK) return st;
  for (int64_t p = 0; p < *nnz_out; ++p) vals[p] = 0.0;
  for (int64_t k = 0; k < nent; ++k) {
    int64_t p = find_in_row(row_ptr, col_ind, rows[k], cols[k]);
    vals[p] = v[k];
  }
  return ORC_OK;
}

/* Compressed_row_matrix.hpp:78-89 */
double orc_csr_at(const int64_t* row_ptr, const int64_t* col_ind,
                  const double* vals, int64_t r, int64_t c) {
  int64_t p = find_in_row(row_ptr, col_ind, r, c);
  return p >= 0 ? vals[p] : 0.0;
}

/* sparse_blas.hpp:17-34: sum starts at T{0}, accumulates in storage order */
void orc_spmv_d(int64_t nrows, const int64_t* row_ptr, const int64_t* col_ind,
                const double* vals, const double* x, double* y) {
  for (int64_t i = 0; i < nrows; ++i) {
    double sum = 0.0;
    for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
      sum += vals[j] * x[col_ind[j]];
    y[i] = sum;
  }
}

void orc_spmv_f(int64_t nrows, const int64_t* row_ptr, const int64_t* col_ind,
                const float* vals, const float* x, float* y) {
  for (int64_t i = 0; i < nrows; ++i) {
    float sum = 0.0f;
    for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
      sum += vals[j] * x[col_ind[j]];
    y[i] = sum;
  }
}

/* unary.hpp:18-39 */
void orc_extract_diagonal(int64_t nrows, int64_t ncols, const int64_t* row_ptr,
                          const int64_t* col_ind, const double* vals,
                          double* d) {
  int64_t len = nrows < ncols ? nrows : ncols;
  for (int64_t i = 0; i < len; ++i) {
    int64_t p = find_in_row(row_ptr, col_ind, i, i);
    d[i] = p >= 0 ? vals[p] : 0.0;
  }
}

/* jacobi.hpp:25-36 */
int orc_jacobi(int64_t nrows, int64_t ncols, const int64_t* row_ptr,
               const int64_t* col_ind, const double* vals, double* inv_diag) {
  int64_t len = nrows < ncols ? nrows : ncols;
  orc_extract_diagonal(nrows, ncols, row_ptr, col_ind, vals, inv_diag);
  for (int64_t i = 0; i < len; ++i) {
    if (inv_diag[i] == 0.0) return ORC_INVALID_ARGUMENT;
    inv_diag[i] = 1.0 / inv_diag[i];
  }
  return ORC_OK;
}

/* jacobi.hpp:42-50 */
void orc_jacobi_apply_d(int64_t n, const double* inv_diag, const double* r,
                        double* z) {
  for (int64_t i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
}

/* ilu.hpp:35-150 */
int orc_ilu0(int64_t nrows, int64_t ncols, const int64_t* rp,
             const int64_t* ci, const double* a_vals, int modified,
             int64_t* l_rp, int64_t* l_ci, double* l_vals, int64_t* u_rp,
             int64_t* u_ci, double* u_vals) {
  int64_t n = nrows;
  if (nrows != ncols) return ORC_INVALID_ARGUMENT; /* ilu.hpp:40-42 */
  int64_t nnz = rp[n];
  double* w = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
  int64_t* diag_pos = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  int64_t* col_map = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  if (!w || !diag_pos || !col_map) {
    free(w);
    free(diag_pos);
    free(col_map);
    return ORC_NO_MEMORY;
  }
  memcpy(w, a_vals, (size_t)nnz * sizeof(double));
  /* step 2: diagonal positions, ilu.hpp:54-66 */
  for (int64_t i = 0; i < n; ++i) {
    int found = 0;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
      if (ci[p] == i) {
        diag_pos[i] = p;
        found = 1;
        break;
      }
    }
    if (!found) {
      free(w);
      free(diag_pos);
      free(col_map);
      return ORC_INVALID_ARGUMENT;
    }
  }
  for (int64_t i = 0; i < n; ++i) col_map[i] = -1;
  /* step 4: IKJ elimination, ilu.hpp:72-115 */
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) col_map[ci[p]] = p;
    double dropped = 0.0;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
      int64_t k = ci[p];
      if (k >= i) break;
      w[p] /= w[diag_pos[k]];
      double l_ik = w[p];
      for (int64_t q = diag_pos[k] + 1; q < rp[k + 1]; ++q) {
        int64_t j = ci[q];
        int64_t pos = col_map[j];
        if (pos >= 0)
          w[pos] -= l_ik * w[q];
        else if (modified)
          dropped += l_ik * w[q];
      }
    }
    if (modified) w[diag_pos[i]] -= dropped;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) col_map[ci[p]] = -1;
  }
  /* step 5: split, ilu.hpp:123-141.  The entries are pushed already in
   * (row, col) order, so the CSR constructor's sort keeps them in place. */
  int64_t lk = 0, uk = 0;
  l_rp[0] = 0;
  u_rp[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
      if (ci[p] < i) {
        l_ci[lk] = ci[p];
        l_vals[lk++] = w[p];
      }
    }
    l_ci[lk] = i;
    l_vals[lk++] = 1.0;
    l_rp[i + 1] = lk;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
      if (ci[p] >= i) {
        u_ci[uk] = ci[p];
        u_vals[uk++] = w[p];
      }
    }
    u_rp[i + 1] = uk;
  }
  free(w);
  free(diag_pos);
  free(col_map);
  return ORC_OK;
}

/* triangular_solve.hpp:27-55 and :67-98, for double and float */
#define ORC_TRISOLVE(SFX, T)                                                   \
  void orc_forward_solve_##SFX(int64_t n, const int64_t* rp,                   \
                               const int64_t* ci, const T* v, const T* b,      \
                               T* x) {                                         \
    if (x != b) memcpy(x, b, (size_t)n * sizeof(T));                           \
    for (int64_t i = 0; i < n; ++i) {                                          \
      for (int64_t p = rp[i]; p < rp[i + 1] - 1; ++p) x[i] -= v[p] * x[ci[p]]; \
      x[i] /= v[rp[i + 1] - 1];                                                \
    }                                                                          \
  }                                                                            \
  void orc_backward_solve_##SFX(int64_t n, const int64_t* rp,                  \
                                const int64_t* ci, const T* v, const T* b,     \
                                T* x) {                                        \
    if (x != b) memcpy(x, b, (size_t)n * sizeof(T));                           \
    for (int64_t ii = n; ii > 0; --ii) {                                       \
      int64_t i = ii - 1;                                                      \
      for (int64_t p = rp[i] + 1; p < rp[i + 1]; ++p) x[i] -= v[p] * x[ci[p]]; \
      x[i] /= v[rp[i]];                                                        \
    }                                                                          \
  }

ORC_TRISOLVE(d, double)
ORC_TRISOLVE(f, float)

/* ------------------------------------------------------------ operators */

#define ORC_OPS(SFX, T, CSR, ILU)                                              \
  void orc_apply_identity_##SFX(const void* ctx, int64_t n, const T* in,       \
                                T* out) {                                      \
    (void)ctx;                                                                 \
    if (out != in) memmove(out, in, (size_t)n * sizeof(T));                    \
  }                                                                            \
  void orc_apply_csr_##SFX(const void* ctx, int64_t n, const T* in, T* out) {  \
    const CSR* A = (const CSR*)ctx;                                            \
    (void)n;                                                                   \
    orc_spmv_##SFX(A->nrows, A->row_ptr, A->col_ind, A->vals, in, out);        \
  }                                                                            \
  void orc_apply_jacobi_##SFX(const void* ctx, int64_t n, const T* in,         \
                              T* out) {                                        \
    const T* inv_d = (const T*)ctx;                                            \
    for (int64_t i = 0; i < n; ++i) out[i] = inv_d[i] * in[i];                 \
  }                                                                            \
  void orc_apply_ilu_##SFX(const void* ctx, int64_t n, const T* in, T* out) {  \
    const ILU* F = (const ILU*)ctx;                                            \
    orc_forward_solve_##SFX(n, F->L.row_ptr, F->L.col_ind, F->L.vals, in,      \
                            F->tmp);                                           \
    orc_backward_solve_##SFX(n, F->U.row_ptr, F->U.col_ind, F->U.vals, F->tmp, \
                             out);                                             \
  }

ORC_OPS(d, double, orc_csr_d, orc_ilu_d)
ORC_OPS(f, float, orc_csr_f, orc_ilu_f)

/* The same operators with the reference's per-call memory traffic, for
 * timing the CPU baseline like the reference runs (results are identical):
 *   the operator lambda over multiply (conjugate_gradient_test.cpp:61-64):
 *   multiply returns a new std::vector<T>(rows, T{0}) (sparse_blas.hpp:25),
 *   which the lambda copies to out;
 *   ilu_apply (ilu.hpp:173-179): forward_solve and backward_solve each
 *   return a new vector initialised from their input (triangular_solve.hpp:
 *   41, 81), and the result is copied to out. */
#define ORC_OPS_REF(SFX, T, CSR, ILU)                                          \
  void orc_apply_csr_ref_##SFX(const void* ctx, int64_t n, const T* in,        \
                               T* out) {                                       \
    const CSR* A = (const CSR*)ctx;                                            \
    (void)n;                                                                   \
    T* y = (T*)malloc((size_t)(A->nrows > 0 ? A->nrows : 1) * sizeof(T));      \
    if (!y) abort();                                                           \
    for (int64_t i = 0; i < A->nrows; ++i) y[i] = (T)0;                        \
    orc_spmv_##SFX(A->nrows, A->row_ptr, A->col_ind, A->vals, in, y);          \
    memcpy(out, y, (size_t)A->nrows * sizeof(T));                              \
    free(y);                                                                   \
  }                                                                            \
  void orc_apply_ilu_ref_##SFX(const void* ctx, int64_t n, const T* in,        \
                               T* out) {                                       \
    const ILU* F = (const ILU*)ctx;                                            \
    const size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(T);                  \
    T* y = (T*)malloc(bytes);                                                  \
    if (!y) abort();                                                           \
    orc_forward_solve_##SFX(n, F->L.row_ptr, F->L.col_ind, F->L.vals, in, y);  \
    T* z = (T*)malloc(bytes);                                                  \
    if (!z) abort();                                                           \
    orc_backward_solve_##SFX(n, F->U.row_ptr, F->U.col_ind, F->U.vals, y, z);  \
    free(y);                                                                   \
    memcpy(out, z, (size_t)n * sizeof(T));                                     \
    free(z);                                                                   \
  }

ORC_OPS_REF(d, double, orc_csr_d, orc_ilu_d)
ORC_OPS_REF(f, float, orc_csr_f, orc_ilu_f)

/* Approximate ilu_apply (no reference counterpart; it replaces the exact
 * triangular solves of ilu.hpp:173-179 / triangular_solve.hpp:27-98 by k
 * Jacobi iterations each):
 *   y_0 = r, y_j = r - L_s y_{j-1};  z_0 = y_k / d, z_j = (y_k - U_s z_{j-1}) / d
 * L_s: every L row but its last entry (the unit diagonal); U_s: every U
 * row but its first (the diagonal d).  Row sums in storage order from 0. */
#define ORC_ILU_JAC(SFX, T, JAC)                                               \
  void orc_apply_ilu_jacobi_##SFX(const void* ctx, int64_t n, const T* in,     \
                                  T* out) {                                    \
    const JAC* F = (const JAC*)ctx;                                            \
    const int64_t k = F->sweeps;                                               \
    T* y = F->tmp;                                                             \
    T* w = F->tmp + n;                                                         \
    T* z = F->tmp + 2 * n;                                                     \
    const int64_t *lrp = F->L.row_ptr, *lci = F->L.col_ind;                    \
    const int64_t *urp = F->U.row_ptr, *uci = F->U.col_ind;                    \
    memcpy(y, in, (size_t)n * sizeof(T));                                      \
    for (int64_t j = 0; j < k; ++j) {                                          \
      for (int64_t i = 0; i < n; ++i) {                                        \
        T t = (T)0;                                                            \
        for (int64_t p = lrp[i]; p < lrp[i + 1] - 1; ++p)                      \
          t += F->L.vals[p] * y[lci[p]];                                       \
        w[i] = in[i] - t;                                                      \
      }                                                                        \
      memcpy(y, w, (size_t)n * sizeof(T));                                     \
    }                                                                          \
    for (int64_t i = 0; i < n; ++i) z[i] = y[i] / F->U.vals[urp[i]];           \
    for (int64_t j = 0; j < k; ++j) {                                          \
      for (int64_t i = 0; i < n; ++i) {                                        \
        T t = (T)0;                                                            \
        for (int64_t p = urp[i] + 1; p < urp[i + 1]; ++p)                      \
          t += F->U.vals[p] * z[uci[p]];                                       \
        w[i] = (y[i] - t) / F->U.vals[urp[i]];                                 \
      }                                                                        \
      memcpy(z, w, (size_t)n * sizeof(T));                                     \
    }                                                                          \
    memcpy(out, z, (size_t)n * sizeof(T));                                     \
  }

ORC_ILU_JAC(d, double, orc_ilu_jacobi_d)
ORC_ILU_JAC(f, float, orc_ilu_jacobi_f)

/* multiply_transpose(A, x): sparse_blas.hpp:36-52 -- y = 0, then for each row
 * i in order and each entry (i, j) in storage order y_j += a_ij x_i */
#define ORC_SPMV_T(SFX, T, CSR)                                                \
  void orc_spmv_transpose_##SFX(int64_t nrows, int64_t ncols,                  \
                                const int64_t* rp, const int64_t* ci,          \
                                const T* v, const T* x, T* y) {                \
    for (int64_t j = 0; j < ncols; ++j) y[j] = (T)0;                           \
    for (int64_t i = 0; i < nrows; ++i)                                        \
      for (int64_t p = rp[i]; p < rp[i + 1]; ++p) y[ci[p]] += v[p] * x[i];     \
  }                                                                            \
  void orc_apply_csr_transpose_##SFX(const void* ctx, int64_t n, const T* in,  \
                                     T* out) {                                 \
    const CSR* A = (const CSR*)ctx;                                            \
    (void)n;                                                                   \
    orc_spmv_transpose_##SFX(A->nrows, A->ncols, A->row_ptr, A->col_ind,       \
                             A->vals, in, out);                                \
  }
ORC_SPMV_T(d, double, orc_csr_d)
ORC_SPMV_T(f, float, orc_csr_f)

double orc_true_relres_d(int64_t n, const int64_t* rp, const int64_t* ci,
                         const double* v, const double* b, const double* x) {
  double rr = 0.0, bb = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * x[ci[p]];
    double r = b[i] - s;
    rr += r * r;
    bb += b[i] * b[i];
  }
  return sqrt(rr) / sqrt(bb);
}

/* ---------------------------------------------------------- generators */

uint64_t orc_splitmix64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_uniform_pm1(uint64_t seed, int64_t n, double* out) {
  for (int64_t k = 0; k < n; ++k) {
    double u = (double)(orc_splitmix64_at(seed, (uint64_t)k) >> 11) *
               (1.0 / 9007199254740992.0);
    out[k] = 2.0 * u - 1.0;
  }
}

#define PUSH(r, c, val)   \
  do {                    \
    if (rows) {           \
      rows[m] = (r);      \
      cols[m] = (c);      \
      v[m] = (val);       \
    }                     \
    ++m;                  \
  } while (0)

/* matgen.hpp:104-117 */
int64_t orc_gen_tridiagonal(int64_t n, double sub, double diag, double super,
                            int64_t* rows, int64_t* cols, double* v) {
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (i > 0) PUSH(i, i - 1, sub);
    PUSH(i, i, diag);
    if (i + 1 < n) PUSH(i, i + 1, super);
  }
  return m;
}

/* matgen.hpp:173-208 (shift == 0) and :222-258 (diagonal degree + shift) */
int64_t orc_gen_poisson_2d(int64_t nx, int64_t ny, double shift,
                           int64_t* rows, int64_t* cols, double* v) {
  int64_t m = 0;
  for (int64_t r = 0; r < ny; ++r) {
    for (int64_t c = 0; c < nx; ++c) {
      int64_t node = r * nx + c;
      int64_t degree = 0;
      if (c > 0) { PUSH(node, node - 1, -1.0); ++degree; }
      if (c + 1 < nx) { PUSH(node, node + 1, -1.0); ++degree; }
      if (r > 0) { PUSH(node, node - nx, -1.0); ++degree; }
      if (r + 1 < ny) { PUSH(node, node + nx, -1.0); ++degree; }
      PUSH(node, node, (double)degree + shift);
    }
  }
  return m;
}

/* matgen.hpp:274-325 (upwind; diagonal = -sum of the present off-diagonals) */
int64_t orc_gen_convection_diffusion_2d(int64_t nx, int64_t ny, double d,
                                        double cx, double cy, int64_t* rows,
                                        int64_t* cols, double* v) {
  int64_t m = 0;
  for (int64_t r = 0; r < ny; ++r) {
    for (int64_t c = 0; c < nx; ++c) {
      int64_t node = r * nx + c;
      double dv = 0.0;
      if (c > 0) { double val = -d - cx; PUSH(node, node - 1, val); dv -= val; }
      if (c + 1 < nx) { double val = -d; PUSH(node, node + 1, val); dv -= val; }
      if (r > 0) { double val = -d - cy; PUSH(node, node - nx, val); dv -= val; }
      if (r + 1 < ny) { double val = -d; PUSH(node, node + nx, val); dv -= val; }
      PUSH(node, node, dv);
    }
  }
  return m;
}

/* matgen.hpp:339-386 */
int64_t orc_gen_convdiff_centered_2d(int64_t nx, int64_t ny, double d,
                                     double cx, double cy, double shift,
                                     int64_t* rows, int64_t* cols, double* v) {
  int64_t m = 0;
  for (int64_t r = 0; r < ny; ++r) {
    for (int64_t c = 0; c < nx; ++c) {
      int64_t node = r * nx + c;
      int64_t degree = 0;
      if (c > 0) { PUSH(node, node - 1, -(d + cx)); ++degree; }
      if (c + 1 < nx) { PUSH(node, node + 1, -(d - cx)); ++degree; }
      if (r > 0) { PUSH(node, node - nx, -(d + cy)); ++degree; }
      if (r + 1 < ny) { PUSH(node, node + nx, -(d - cy)); ++degree; }
      PUSH(node, node, (double)degree * d + shift);
    }
  }
  return m;
}

/* matgen.hpp:147-164 */
int64_t orc_gen_nonsymmetric_sample(int64_t* rows, int64_t* cols, double* v) {
  int64_t m = 0;
  PUSH(0, 0, 4.0); PUSH(0, 1, -1.0); PUSH(0, 2, 0.5);
  PUSH(1, 0, -0.5); PUSH(1, 1, 4.0); PUSH(1, 3, -1.0);
  PUSH(2, 0, 0.3); PUSH(2, 2, 4.0); PUSH(2, 3, -0.8);
  PUSH(3, 1, -0.6); PUSH(3, 2, 0.2); PUSH(3, 3, 4.0);
  return m;
}

#undef PUSH

/* 7-point stencils written straight into CSR.  Columns of a row ascend:
 * -nx*ny, -nx, -1, 0, +1, +nx, +nx*ny. */
#define EMIT(col, val)            \
  do {                            \
    if (col_ind) {                \
      col_ind[m] = (col);         \
      vals[m] = (val);            \
    }                             \
    ++m;                          \
  } while (0)

int64_t orc_gen_poisson_3d_csr(int64_t nx, int64_t ny, int64_t nz,
                               int64_t* row_ptr, int64_t* col_ind,
                               double* vals) {
  int64_t m = 0, plane = nx * ny;
  if (row_ptr) row_ptr[0] = 0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t row = k * plane + j * nx + i;
        if (k > 0) EMIT(row - plane, -1.0);
        if (j > 0) EMIT(row - nx, -1.0);
        if (i > 0) EMIT(row - 1, -1.0);
        EMIT(row, 6.0);
        if (i + 1 < nx) EMIT(row + 1, -1.0);
        if (j + 1 < ny) EMIT(row + nx, -1.0);
        if (k + 1 < nz) EMIT(row + plane, -1.0);
        if (row_ptr) row_ptr[row + 1] = m;
      }
  return m;
}

int64_t orc_gen_convdiff_3d_csr(int64_t nx, int64_t ny, int64_t nz, double d,
                                double cx, double cy, double cz,
                                int64_t* row_ptr, int64_t* col_ind,
                                double* vals) {
  int64_t m = 0, plane = nx * ny;
  double diag = 6.0 * d + cx + cy + cz;
  if (row_ptr) row_ptr[0] = 0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t row = k * plane + j * nx + i;
        if (k > 0) EMIT(row - plane, -d - cz);
        if (j > 0) EMIT(row - nx, -d - cy);
        if (i > 0) EMIT(row - 1, -d - cx);
        EMIT(row, diag);
        if (i + 1 < nx) EMIT(row + 1, -d);
        if (j + 1 < ny) EMIT(row + nx, -d);
        if (k + 1 < nz) EMIT(row + plane, -d);
        if (row_ptr) row_ptr[row + 1] = m;
      }
  return m;
}

#undef EMIT

/* ------------------------------------------------------------ permutations
 * perm maps old index -> new index.  permutation.cpp:15-35 (validity,
 * inverse), permutation.hpp:41-148 (rperm: row r of the result is old row
 * inv[r]; cperm: column c becomes perm[c], each row re-sorted by the new
 * columns; dperm: both, P A P^T).  Values travel with their entries. */
int orc_is_valid_permutation(int64_t n, const int64_t* perm) {
  char* seen = (char*)calloc((size_t)(n > 0 ? n : 1), 1);
  int ok = 1;
  for (int64_t k = 0; k < n && ok; ++k) {
    if (perm[k] < 0 || perm[k] >= n || seen[perm[k]]) ok = 0;
    else seen[perm[k]] = 1;
  }
  free(seen);
  return ok;
}

typedef struct {
  int64_t c;
  double v;
} orc_cv;

static int cmp_cv(const void* a, const void* b) {
  int64_t x = ((const orc_cv*)a)->c, y = ((const orc_cv*)b)->c;
  return (x > y) - (x < y);
}

/* mode: 1 rows, 2 columns, 3 both.  Output arrays: row_ptr n+1, col_ind
 * and vals nnz (vals may be NULL for a pattern). */
int orc_csr_permute(int64_t nrows, int64_t ncols, const int64_t* rp,
                    const int64_t* ci, const double* v, const int64_t* perm,
                    int mode, int64_t* orp, int64_t* oci, double* ov) {
  const int64_t np = (mode & 1) ? nrows : ncols;
  if ((mode & 1) && (mode & 2) && nrows != ncols) return ORC_INVALID_ARGUMENT;
  if (!orc_is_valid_permutation(np, perm)) return ORC_INVALID_ARGUMENT;
  int64_t* inv = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrows > 0 ? nrows : 1));
  for (int64_t r = 0; r < nrows; ++r) inv[r] = r;
  if (mode & 1)
    for (int64_t k = 0; k < nrows; ++k) inv[perm[k]] = k;
  int64_t maxlen = 0;
  for (int64_t r = 0; r < nrows; ++r)
    if (rp[r + 1] - rp[r] > maxlen) maxlen = rp[r + 1] - rp[r];
  orc_cv* row = (orc_cv*)malloc(sizeof(orc_cv) * (size_t)(maxlen > 0 ? maxlen : 1));
  int64_t out = 0;
  orp[0] = 0;
  for (int64_t nr = 0; nr < nrows; ++nr) {
    const int64_t o = inv[nr];
    int64_t len = 0;
    for (int64_t p = rp[o]; p < rp[o + 1]; ++p, ++len) {
      row[len].c = (mode & 2) ? perm[ci[p]] : ci[p];
      row[len].v = v ? v[p] : 0.0;
    }
    if (mode & 2) qsort(row, (size_t)len, sizeof(orc_cv), cmp_cv);
    for (int64_t k = 0; k < len; ++k, ++out) {
      oci[out] = row[k].c;
      if (ov) ov[out] = row[k].v;
    }
    orp[nr + 1] = out;
  }
  free(row);
  free(inv);
  return ORC_OK;
}
