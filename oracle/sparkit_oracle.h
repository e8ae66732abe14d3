/*
 * sparkit_oracle.h -- CPU restatement of sparkit's CSR SpMV + Krylov solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP implementation in sparkit_amd/.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  Product code never
 * links, imports or calls anything in oracle/.
 *
 * Every function restates one reference algorithm (cited file:line, paths
 * relative to the sparkit source tree) in plain C with the same operation
 * order, so that integer outputs are bit-exact and floating-point outputs
 * match the reference's sequential left-to-right arithmetic.  The library is
 * compiled with -ffp-contract=off so that no multiply-add is fused (the
 * reference's x86-64 build has no FMA either).
 *
 * Index width is int64, as in the reference (size_type = std::ptrdiff_t,
 * sparkit/data/import.hpp:19).
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_csr_pattern     reference KATs compressed_data_test.cpp:142-203 +
 *                       brute force (python set/sort) + stencil nnz closed forms
 *   orc_make_matrix     matgen_test.cpp:49-72 round trip + last-wins rule
 *   orc_spmv_*          sparse_blas_test.cpp:36-95 KATs + exact integer
 *                       arithmetic + scipy csr_matrix @ x
 *   orc_extract_diagonal / orc_jacobi*   unary_test.cpp:35-95, jacobi_test.cpp:50-95
 *   orc_ilu0            ilu_test.cpp:54-107,288-324 (diagonal, tridiagonal exact LU,
 *                       MILU row sums) + dense IKJ on small matrices with fill
 *   orc_forward/backward_solve  triangular_solve_test.cpp KATs + dense substitution
 *   orc_cg/gmres/fgmres/bicgstab  solver tests (*_test.cpp): solution recovery,
 *                       zero rhs, max_iterations, residual monotonicity,
 *                       CG exact in <= n steps, GMRES(m>=n) exact at step n,
 *                       scipy spsolve reference solutions
 *   generators          matgen_test.cpp KATs + closed forms (nnz, row sums,
 *                       symmetry)
 */
#ifndef SPARKIT_ORACLE_H
#define SPARKIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (shared meaning with the HIP library's C ABI). */
enum {
  ORC_OK = 0,
  ORC_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
  ORC_LOGIC_ERROR = 2,      /* std::logic_error (bad Index / Shape) */
  ORC_NO_MEMORY = 3,
  ORC_BREAKDOWN = 5         /* numerical breakdown the reference does not detect */
};

/* ------------------------------------------------------------------ CSR */

/* Compressed_row_sparsity(Shape, indices): sort (row, col), unique, counts,
 * prefix sum.  Compressed_row_sparsity_Impl.cpp:10-40, Index.cpp:5-9,
 * Shape.cpp:5-11.  col_ind must hold nidx entries; *nnz_out receives the
 * number of unique positions. */
int orc_csr_pattern(int64_t nrows, int64_t ncols, int64_t nidx,
                    const int64_t* rows, const int64_t* cols,
                    int64_t* row_ptr, int64_t* col_ind, int64_t* nnz_out);

/* make_matrix(shape, entries): pattern from the entries, then each entry's
 * value placed by a scan of its row; a later duplicate overwrites an earlier
 * one.  matgen.hpp:25-52. */
int orc_make_matrix(int64_t nrows, int64_t ncols, int64_t nent,
                    const int64_t* rows, const int64_t* cols, const double* v,
                    int64_t* row_ptr, int64_t* col_ind, double* vals,
                    int64_t* nnz_out);

/* Compressed_row_matrix::operator()(r, c): binary search, 0 if absent.
 * Compressed_row_matrix.hpp:78-89. */
double orc_csr_at(const int64_t* row_ptr, const int64_t* col_ind,
                  const double* vals, int64_t r, int64_t c);

/* Permutations (perm[old] = new): permutation.cpp:15-35 and
 * permutation.hpp:41-148.  mode 1 = rperm, 2 = cperm, 3 = dperm (P A P^T).
 * ORC_INVALID_ARGUMENT for an invalid permutation (or dperm of a
 * non-square matrix). */
int orc_is_valid_permutation(int64_t n, const int64_t* perm);
int orc_csr_permute(int64_t nrows, int64_t ncols, const int64_t* rp,
                    const int64_t* ci, const double* v, const int64_t* perm,
                    int mode, int64_t* orp, int64_t* oci, double* ov);

/* multiply(A, x): y_i = sum_j a_ij x_cj in storage order from 0.
 * sparse_blas.hpp:17-34. */
void orc_spmv_d(int64_t nrows, const int64_t* row_ptr, const int64_t* col_ind,
                const double* vals, const double* x, double* y);
void orc_spmv_f(int64_t nrows, const int64_t* row_ptr, const int64_t* col_ind,
                const float* vals, const float* x, float* y);

/* extract_diagonal(A): unary.hpp:18-39.  d has min(nrows, ncols) entries. */
void orc_extract_diagonal(int64_t nrows, int64_t ncols, const int64_t* row_ptr,
                          const int64_t* col_ind, const double* vals,
                          double* d);

/* jacobi(A): inverse diagonal, ORC_INVALID_ARGUMENT on an exact zero.
 * jacobi.hpp:25-36. */
int orc_jacobi(int64_t nrows, int64_t ncols, const int64_t* row_ptr,
               const int64_t* col_ind, const double* vals, double* inv_diag);

/* jacobi_apply: z_i = inv_d_i * r_i.  jacobi.hpp:42-50. */
void orc_jacobi_apply_d(int64_t n, const double* inv_diag, const double* r,
                        double* z);

/* ilu0_impl(A, modified): IKJ ILU(0)/MILU(0) on A's pattern, then split
 * into L (strictly lower + unit diagonal last) and U (diagonal first +
 * upper).  ilu.hpp:35-168.  L arrays need nnz_lower + n entries, U arrays
 * nnz_upper + n (nnz_upper counts the diagonal's absence: use nnz(A) as a
 * safe capacity for both).  Returns ORC_INVALID_ARGUMENT for a non-square
 * matrix or a missing diagonal (ilu.hpp:40-41, 63-65). */
int orc_ilu0(int64_t nrows, int64_t ncols, const int64_t* row_ptr,
             const int64_t* col_ind, const double* vals, int modified,
             int64_t* l_row_ptr, int64_t* l_col_ind, double* l_vals,
             int64_t* u_row_ptr, int64_t* u_col_ind, double* u_vals);

/* forward_solve(L, b): row order, diagonal is the LAST entry of each row.
 * triangular_solve.hpp:27-55.  backward_solve(U, b): reverse row order,
 * diagonal is the FIRST entry of each row.  triangular_solve.hpp:67-98. */
void orc_forward_solve_d(int64_t n, const int64_t* rp, const int64_t* ci,
                         const double* v, const double* b, double* x);
void orc_backward_solve_d(int64_t n, const int64_t* rp, const int64_t* ci,
                          const double* v, const double* b, double* x);
void orc_forward_solve_f(int64_t n, const int64_t* rp, const int64_t* ci,
                         const float* v, const float* b, float* x);
void orc_backward_solve_f(int64_t n, const int64_t* rp, const int64_t* ci,
                          const float* v, const float* b, float* x);

/* ------------------------------------------------------------ operators */

/* The reference's callable concept op(first, last, out)
 * (conjugate_gradient.hpp:103-109) restated as a C callback. */
typedef void (*orc_op_d)(const void* ctx, int64_t n, const double* in,
                         double* out);
typedef void (*orc_op_f)(const void* ctx, int64_t n, const float* in,
                         float* out);

typedef struct {
  int64_t nrows, ncols;
  const int64_t* row_ptr;
  const int64_t* col_ind;
  const double* vals;
} orc_csr_d;

typedef struct {
  int64_t nrows, ncols;
  const int64_t* row_ptr;
  const int64_t* col_ind;
  const float* vals;
} orc_csr_f;

typedef struct {
  orc_csr_d L, U;
  double* tmp; /* n scratch entries */
} orc_ilu_d;

typedef struct {
  orc_csr_f L, U;
  float* tmp;
} orc_ilu_f;

/* approximate ILU apply: factors, Jacobi sweeps per triangle, 3n scratch */
typedef struct {
  orc_csr_d L, U;
  int64_t sweeps;
  double* tmp;
} orc_ilu_jacobi_d;

typedef struct {
  orc_csr_f L, U;
  int64_t sweeps;
  float* tmp;
} orc_ilu_jacobi_f;

/* identity slot: conjugate_gradient_test.cpp:45-47 */
void orc_apply_identity_d(const void* ctx, int64_t n, const double* in, double* out);
void orc_apply_identity_f(const void* ctx, int64_t n, const float* in, float* out);
/* operator lambda over multiply(A, x): conjugate_gradient_test.cpp:61-64 */
void orc_apply_csr_d(const void* ctx, int64_t n, const double* in, double* out);
void orc_apply_csr_f(const void* ctx, int64_t n, const float* in, float* out);
/* the same two operators with the reference's per-call allocation and
 * copy (sparse_blas.hpp:25, ilu.hpp:173-179, triangular_solve.hpp:41,81);
 * identical results, used to time the CPU baseline */
void orc_apply_csr_ref_d(const void* ctx, int64_t n, const double* in, double* out);
void orc_apply_csr_ref_f(const void* ctx, int64_t n, const float* in, float* out);
void orc_apply_ilu_ref_d(const void* ctx, int64_t n, const double* in, double* out);
void orc_apply_ilu_ref_f(const void* ctx, int64_t n, const float* in, float* out);
/* multiply_transpose(A, x): sparse_blas.hpp:36-52 */
void orc_spmv_transpose_d(int64_t nrows, int64_t ncols, const int64_t* rp,
                          const int64_t* ci, const double* v, const double* x,
                          double* y);
void orc_spmv_transpose_f(int64_t nrows, int64_t ncols, const int64_t* rp,
                          const int64_t* ci, const float* v, const float* x,
                          float* y);
void orc_apply_csr_transpose_d(const void* ctx, int64_t n, const double* in,
                               double* out);
void orc_apply_csr_transpose_f(const void* ctx, int64_t n, const float* in,
                               float* out);
/* jacobi_apply lambda: jacobi_test.cpp:118-121; ctx points at inv_diag */
void orc_apply_jacobi_d(const void* ctx, int64_t n, const double* in, double* out);
void orc_apply_jacobi_f(const void* ctx, int64_t n, const float* in, float* out);
/* ilu_apply lambda: ilu.hpp:173-179, ilu_test.cpp:127-130 */
void orc_apply_ilu_d(const void* ctx, int64_t n, const double* in, double* out);
void orc_apply_ilu_f(const void* ctx, int64_t n, const float* in, float* out);
/* approximate ilu_apply by Jacobi sweeps (no reference counterpart; ctx
 * points at an orc_ilu_jacobi_*) */
void orc_apply_ilu_jacobi_d(const void* ctx, int64_t n, const double* in,
                            double* out);
void orc_apply_ilu_jacobi_f(const void* ctx, int64_t n, const float* in,
                            float* out);

/* ------------------------------------------------------------- solvers */

/* Config and summary mirror CGConfig / Gmres_config / Fgmres_config /
 * Bicgstab_config and the *Summary structs (conjugate_gradient.hpp:26-74,
 * gmres.hpp:24-68, fgmres.hpp:24-68, bicgstab.hpp:24-63).  restart is
 * restart_iterations for CG and restart_dimension for (F)GMRES; BiCGSTAB
 * ignores it.  history (may be NULL) receives one residual per entry the
 * reference would push into iteration_residuals; it must hold
 * max_iterations entries. */
typedef struct {
  double tolerance;
  int64_t restart;
  int64_t max_iterations;
} orc_config_d;

typedef struct {
  float tolerance;
  int64_t restart;
  int64_t max_iterations;
} orc_config_f;

typedef struct {
  double residual_norm;
  int64_t computed_iterations;
  int32_t converged;
  int32_t status; /* ORC_OK or ORC_BREAKDOWN / ORC_INVALID_ARGUMENT */
  int64_t history_len;
} orc_summary_d;

typedef struct {
  float residual_norm;
  int64_t computed_iterations;
  int32_t converged;
  int32_t status;
  int64_t history_len;
} orc_summary_f;

/* ConjugateGradientSolver: conjugate_gradient.hpp:122-312 */
void orc_cg_d(int64_t n, const double* b, double* x, const orc_config_d* cfg,
              orc_op_d A, const void* actx, orc_op_d ML, const void* mlctx,
              orc_op_d MR, const void* mrctx, orc_summary_d* out,
              double* history);
void orc_cg_f(int64_t n, const float* b, float* x, const orc_config_f* cfg,
              orc_op_f A, const void* actx, orc_op_f ML, const void* mlctx,
              orc_op_f MR, const void* mrctx, orc_summary_f* out,
              float* history);

/* GmresSolver: gmres.hpp:99-375 */
void orc_gmres_d(int64_t n, const double* b, double* x, const orc_config_d* cfg,
                 orc_op_d A, const void* actx, orc_op_d ML, const void* mlctx,
                 orc_op_d MR, const void* mrctx, orc_summary_d* out,
                 double* history);
void orc_gmres_f(int64_t n, const float* b, float* x, const orc_config_f* cfg,
                 orc_op_f A, const void* actx, orc_op_f ML, const void* mlctx,
                 orc_op_f MR, const void* mrctx, orc_summary_f* out,
                 float* history);

/* FgmresSolver: fgmres.hpp:96-428 (single right preconditioner slot) */
void orc_fgmres_d(int64_t n, const double* b, double* x,
                  const orc_config_d* cfg, orc_op_d A, const void* actx,
                  orc_op_d M, const void* mctx, orc_summary_d* out,
                  double* history);
void orc_fgmres_f(int64_t n, const float* b, float* x, const orc_config_f* cfg,
                  orc_op_f A, const void* actx, orc_op_f M, const void* mctx,
                  orc_summary_f* out, float* history);

/* BicgstabSolver: bicgstab.hpp:86-361 */
void orc_bicgstab_d(int64_t n, const double* b, double* x,
                    const orc_config_d* cfg, orc_op_d A, const void* actx,
                    orc_op_d ML, const void* mlctx, orc_op_d MR,
                    const void* mrctx, orc_summary_d* out, double* history);
void orc_bicgstab_f(int64_t n, const float* b, float* x,
                    const orc_config_f* cfg, orc_op_f A, const void* actx,
                    orc_op_f ML, const void* mlctx, orc_op_f MR,
                    const void* mrctx, orc_summary_f* out, float* history);

/* CgsSolver: cgs.hpp:86-395 (two matvecs per step, counted) */
void orc_cgs_d(int64_t n, const double* b, double* x, const orc_config_d* cfg,
               orc_op_d A, const void* actx, orc_op_d ML, const void* mlctx,
               orc_op_d MR, const void* mrctx, orc_summary_d* out,
               double* history);
void orc_cgs_f(int64_t n, const float* b, float* x, const orc_config_f* cfg,
               orc_op_f A, const void* actx, orc_op_f ML, const void* mlctx,
               orc_op_f MR, const void* mrctx, orc_summary_f* out,
               float* history);

/* MinresSolver: minres.hpp:80-395 (M = M_L M_R, one matvec per step) */
void orc_minres_d(int64_t n, const double* b, double* x,
                  const orc_config_d* cfg, orc_op_d A, const void* actx,
                  orc_op_d ML, const void* mlctx, orc_op_d MR,
                  const void* mrctx, orc_summary_d* out, double* history);
void orc_minres_f(int64_t n, const float* b, float* x, const orc_config_f* cfg,
                  orc_op_f A, const void* actx, orc_op_f ML, const void* mlctx,
                  orc_op_f MR, const void* mrctx, orc_summary_f* out,
                  float* history);

/* QmrSolver: qmr.hpp:86-491; At is the transpose operator */
void orc_qmr_d(int64_t n, const double* b, double* x, const orc_config_d* cfg,
               orc_op_d A, const void* actx, orc_op_d At, const void* atctx,
               orc_op_d ML, const void* mlctx, orc_op_d MR, const void* mrctx,
               orc_summary_d* out, double* history);
void orc_qmr_f(int64_t n, const float* b, float* x, const orc_config_f* cfg,
               orc_op_f A, const void* actx, orc_op_f At, const void* atctx,
               orc_op_f ML, const void* mlctx, orc_op_f MR, const void* mrctx,
               orc_summary_f* out, float* history);

/* BicgSolver: bicg.hpp:86-397; At is the transpose operator */
void orc_bicg_d(int64_t n, const double* b, double* x, const orc_config_d* cfg,
                orc_op_d A, const void* actx, orc_op_d At, const void* atctx,
                orc_op_d ML, const void* mlctx, orc_op_d MR, const void* mrctx,
                orc_summary_d* out, double* history);
void orc_bicg_f(int64_t n, const float* b, float* x, const orc_config_f* cfg,
                orc_op_f A, const void* actx, orc_op_f At, const void* atctx,
                orc_op_f ML, const void* mlctx, orc_op_f MR, const void* mrctx,
                orc_summary_f* out, float* history);

/* ||b - A x||_2 / ||b||_2 in double, sequential sums. */
double orc_true_relres_d(int64_t n, const int64_t* rp, const int64_t* ci,
                         const double* v, const double* b, const double* x);

/* ---------------------------------------------------------- generators */

/* splitmix64 (Steele, Lea & Flood 2014): state += 0x9E3779B97F4A7C15 then
 * the standard mix.  u01 = (z >> 11) * 2^-53; uniform[-1,1] = 2 u01 - 1.
 * Element k of stream `seed` is mix(seed + (k + 1) * gamma). */
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t k);
void orc_uniform_pm1(uint64_t seed, int64_t n, double* out);

/* Entry lists of the reference generators, in the reference's push order
 * (fed to orc_make_matrix).  Each returns the entry count; pass NULL
 * buffers to query it.
 *   tridiagonal_matrix  matgen.hpp:104-117
 *   poisson_2d          matgen.hpp:173-215 (shift = 0 gives the singular
 *                       Neumann Laplacian; shift != 0 is :222-265)
 *   convection_diffusion_2d (upwind)  matgen.hpp:274-325
 *   convdiff_centered_2d              matgen.hpp:339-386
 *   nonsymmetric_sample               matgen.hpp:147-164 */
int64_t orc_gen_tridiagonal(int64_t n, double sub, double diag, double super,
                            int64_t* rows, int64_t* cols, double* v);
int64_t orc_gen_poisson_2d(int64_t nx, int64_t ny, double shift,
                           int64_t* rows, int64_t* cols, double* v);
int64_t orc_gen_convection_diffusion_2d(int64_t nx, int64_t ny, double d,
                                        double cx, double cy, int64_t* rows,
                                        int64_t* cols, double* v);
int64_t orc_gen_convdiff_centered_2d(int64_t nx, int64_t ny, double d,
                                     double cx, double cy, double shift,
                                     int64_t* rows, int64_t* cols, double* v);
int64_t orc_gen_nonsymmetric_sample(int64_t* rows, int64_t* cols, double* v);

/* Generators with no reference counterpart, written directly in CSR
 * (row-major natural ordering, columns ascending).  Pass NULL arrays to
 * query nnz through the return value.
 *   poisson_3d_dirichlet: 7-point, diagonal 6, off-diagonal -1; a missing
 *     neighbour (boundary) is dropped, the diagonal stays 6.
 *   convdiff_3d_upwind: 7-point first-order upwind extending
 *     convection_diffusion_2d's stencil (matgen.hpp:292-321) to z:
 *     upwind neighbour -(d + c_k), downwind -d, diagonal 6 d + sum_k c_k
 *     regardless of boundary (Dirichlet), so the matrix is nonsingular. */
int64_t orc_gen_poisson_3d_csr(int64_t nx, int64_t ny, int64_t nz,
                               int64_t* row_ptr, int64_t* col_ind,
                               double* vals);
int64_t orc_gen_convdiff_3d_csr(int64_t nx, int64_t ny, int64_t nz, double d,
                                double cx, double cy, double cz,
                                int64_t* row_ptr, int64_t* col_ind,
                                double* vals);

/* Power-law SPD matrix (definition in oracle_gen.c's header comment).
 * orc_powerlaw_cdf fills the 4094-entry Zipf(1.5) table for lengths 3..4096;
 * orc_powerlaw_length draws row i's target length; orc_gen_powerlaw_spd
 * writes the CSR arrays (pass NULL arrays to get nnz). */
void orc_powerlaw_cdf(double* cdf);
int64_t orc_powerlaw_length(const double* cdf, uint64_t seed, int64_t i);
int64_t orc_gen_powerlaw_spd(int64_t n, uint64_t seed, int64_t* row_ptr,
                             int64_t* col_ind, double* vals);

/* 2D anisotropic variable-coefficient diffusion, Q1 finite elements, N x N
 * interior nodes (definition in oracle_gen.c).  orc_aniso_q1_params fills
 * the (N+1)^2 cell coefficients and the 4x4 element table; iso = 1 forces
 * eps = 1, theta = 0.  Arrays hold N^2 + 1 and (3N - 2)^2 entries. */
void orc_aniso_q1_params(int64_t N, uint64_t seed, double sigma, int iso,
                         double* kappa, double* E);
int64_t orc_gen_aniso_q1_2d(int64_t N, uint64_t seed, double sigma, int iso,
                            int64_t* row_ptr, int64_t* col_ind, double* vals);

/* Batched 2D variable-coefficient diffusion, system k (definition in
 * oracle_gen.c): side s_k, cell conductivities (s_k^2) and the s_k^2-row
 * CSR (at most 5 s_k^2 nonzeros). */
int64_t orc_batched_side(uint64_t seed, int64_t k, int64_t smin, int64_t smax);
void orc_batched_kappa(uint64_t seed, int64_t k, int64_t s, double* kappa);
int64_t orc_gen_batched_diffusion(uint64_t seed, int64_t k, int64_t smin,
                                  int64_t smax, int64_t* row_ptr,
                                  int64_t* col_ind, double* vals);

/* ------------------------------------------------- learned SPAI (oracle_spai.c)
 * Features, forward, ||I - MA||_F^2 / n loss and its gradient, and M's
 * values on A's pattern.  _bf16: the device arithmetic (bf16 GEMM operands,
 * fp32 accumulation); _exact: float64 throughout (finite-difference
 * checks).  params: the 6345 canonical parameters; grad may be NULL.  The
 * loss functions return NaN for a matrix the slot rules reject. */
int orc_spai_features(const orc_csr_d* A, float* X /* n x 24 */);
float orc_bf16_round(float x);
double orc_spai_loss_grad_bf16(const float* params, const orc_csr_d* A,
                               double* grad);
double orc_spai_loss_grad_exact(const double* params, const orc_csr_d* A,
                                double* grad);
int orc_spai_predict_bf16(const float* params, const orc_csr_d* A,
                          double* mvals);
/* the symmetric objective (1/n) ||I - S A||_F^2, S = (M + M^T) / 2, and its
 * gradient; NaN also for a pattern that is not structurally symmetric */
double orc_spai_sym_loss_grad_bf16(const float* params, const orc_csr_d* A,
                                   double* grad);
double orc_spai_sym_loss_grad_exact(const double* params, const orc_csr_d* A,
                                    double* grad);
int orc_spai_predict_exact(const double* params, const orc_csr_d* A,
                           double* mvals);

#ifdef __cplusplus
}
#endif

#endif /* SPARKIT_ORACLE_H */
