/*
 * sparkit_hip.h -- C ABI of the MI355X (gfx950) implementation of sparkit's
 * CSR SpMV + preconditioned Krylov solve path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * Vectors passed to the compute entry points are DEVICE pointers (allocate
 * them with spk_malloc, or any hipMalloc'd memory); the *_create entry
 * points take HOST arrays in the reference's layout (int64 indices) and
 * upload them once.  All work is ordered on one library stream
 * (spk_set_stream / spk_get_stream).
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the sparkit source tree, namespace sparkit::data::detail).
 * The reference has no C ABI of its own; its "operator API" is the C++
 * callable concept op(first, last, out) (conjugate_gradient.hpp:103-109).
 * include/sparkit/hip/ wraps this ABI back into that concept.
 */
#ifndef SPARKIT_HIP_H
#define SPARKIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int spk_status;

/* Status codes.  The C++ wrapper maps them to the reference's exception
 * types: INVALID_ARGUMENT -> std::invalid_argument (jacobi.hpp:31,
 * ilu.hpp:41,64, triangular_solve.hpp:33,73), LOGIC_ERROR ->
 * std::logic_error (Index.cpp:8, Shape.cpp:8-10). */
enum {
  SPK_OK = 0,
  SPK_INVALID_ARGUMENT = 1,
  SPK_LOGIC_ERROR = 2,
  SPK_OUT_OF_MEMORY = 3,
  SPK_HIP_ERROR = 4,
  SPK_BREAKDOWN = 5, /* numerical breakdown the reference leaves undetected */
  SPK_NO_DEVICE = 6
};

/* Arithmetic type of a store, vector or solve. */
enum { SPK_F64 = 0, SPK_F32 = 1 };

typedef struct spk_csr_s* spk_csr;         /* device compressed_row store */
typedef struct spk_precond_s* spk_precond; /* device preconditioner */

/* ------------------------------------------------------------- runtime */

/* Message of the last failing call on this thread. */
const char* spk_last_error(void);
/* Select the device and create the library stream (default priority). */
spk_status spk_init(int device);
spk_status spk_finalize(void);
spk_status spk_set_stream(void* hip_stream); /* NULL: library stream */
void* spk_get_stream(void);
spk_status spk_synchronize(void);
spk_status spk_device_info(int* num_cus, char* name, size_t name_len);

spk_status spk_malloc(void** dptr, size_t bytes);
spk_status spk_free(void* dptr);
spk_status spk_memcpy_htod(void* dst, const void* src, size_t bytes);
spk_status spk_memcpy_dtoh(void* dst, const void* src, size_t bytes);
spk_status spk_memcpy_dtod(void* dst, const void* src, size_t bytes);
spk_status spk_memset(void* dst, int value, size_t bytes);

/* Event timer on the library stream: spk_timer_stop returns the elapsed
 * milliseconds between the two marks (synchronises). */
spk_status spk_timer_start(void);
spk_status spk_timer_stop(float* ms);

/* --------------------------------------------------------- CSR store */

/* Replaces Compressed_row_matrix<T>(Compressed_row_sparsity, std::vector<T>)
 * (Compressed_row_matrix.hpp:25-28) as the thing the operator slot reads:
 * uploads row_ptr/col_ind (int64, as Compressed_row_sparsity.hpp:97-107
 * exposes them) narrowed to int32 and the values (double), and builds the
 * row-binned SpMV plan.  Fails with SPK_INVALID_ARGUMENT if nnz or nrows do
 * not fit in int32 or the pattern is malformed (row_ptr not monotone,
 * column out of range). */
spk_status spk_csr_create(int64_t nrows, int64_t ncols, int64_t nnz,
                          const int64_t* row_ptr, const int64_t* col_ind,
                          const double* vals, spk_csr* out);
/* Same from int32 device arrays already in HBM (copied into the store). */
spk_status spk_csr_create_device(int64_t nrows, int64_t ncols, int64_t nnz,
                                 const int32_t* d_row_ptr,
                                 const int32_t* d_col_ind,
                                 const double* d_vals, spk_csr* out);
spk_status spk_csr_destroy(spk_csr A);
spk_status spk_csr_shape(spk_csr A, int64_t* nrows, int64_t* ncols,
                         int64_t* nnz);
/* Copy the store back in the reference layout (int64 indices, double
 * values); any pointer may be NULL. */
spk_status spk_csr_download(spk_csr A, int64_t* row_ptr, int64_t* col_ind,
                            double* vals);
/* Device pointers of the store (int32 row_ptr/col_ind, values of `dtype`;
 * an FP32 copy is made on first request). */
spk_status spk_csr_device_arrays(spk_csr A, int dtype, const int32_t** rp,
                                 const int32_t** ci, const void** vals);
/* Number of row blocks in the SpMV plan and how many are long-row blocks. */
spk_status spk_csr_plan_info(spk_csr A, int64_t* nblocks,
                             int64_t* nlong_blocks);

/* multiply(A, x) (sparse_blas.hpp:17-34): y = A x on device vectors of
 * `dtype` (SPK_F64, or SPK_F32 using the FP32 copy of the values). */
spk_status spk_csr_spmv(spk_csr A, int dtype, const void* d_x, void* d_y);

/* extract_diagonal(A) (unary.hpp:18-39): d has min(nrows, ncols) doubles. */
spk_status spk_csr_extract_diagonal(spk_csr A, double* d_diag);

/* -------------------------------------------------------------- BLAS-1 */

/* std::inner_product over device vectors (the reductions inside the
 * solvers, e.g. conjugate_gradient.hpp:194,236): result written to *host.
 * FP32 vectors are accumulated in double. */
spk_status spk_dot(int dtype, int64_t n, const void* d_x, const void* d_y,
                   double* result);
/* y = a x + b y */
spk_status spk_axpby(int dtype, int64_t n, double a, const void* d_x,
                     double b, void* d_y);
/* dst (dtype_dst) = src (dtype_src) with rounding to nearest */
spk_status spk_convert(int64_t n, int dtype_src, const void* d_src,
                       int dtype_dst, void* d_dst);

/* ------------------------------------------------------ preconditioners */

/* jacobi(A) (jacobi.hpp:25-36): inverse diagonal on the device;
 * SPK_INVALID_ARGUMENT on an exact zero diagonal entry. */
spk_status spk_precond_create_jacobi(spk_csr A, spk_precond* out);
/* ilu0(A) / milu0(A) (ilu.hpp:35-168): factors on A's pattern, computed on
 * the device by level-scheduled IKJ rows; L has its unit diagonal last,
 * U its diagonal first (ilu.hpp:123-141).  SPK_INVALID_ARGUMENT for a
 * non-square matrix or a missing diagonal. */
spk_status spk_precond_create_ilu0(spk_csr A, int modified, spk_precond* out);
/* Identity slot (conjugate_gradient_test.cpp:45-47). */
spk_status spk_precond_create_identity(int64_t n, spk_precond* out);
/* z = M r for an explicit sparse approximate inverse M (a store). */
spk_status spk_precond_create_spmv(spk_csr M, spk_precond* out);
spk_status spk_precond_destroy(spk_precond M);
/* jacobi_apply (jacobi.hpp:42-50) / ilu_apply (ilu.hpp:173-179) / identity
 * copy / M r, on device vectors of `dtype`. */
spk_status spk_precond_apply(spk_precond M, int dtype, const void* d_in,
                             void* d_out);
/* Copy ILU(0) factors back (reference layout); sizes via *_nnz first. */
spk_status spk_ilu0_factor_nnz(spk_precond M, int64_t* nnz_l, int64_t* nnz_u);
spk_status spk_ilu0_download(spk_precond M, int64_t* l_row_ptr,
                             int64_t* l_col_ind, double* l_vals,
                             int64_t* u_row_ptr, int64_t* u_col_ind,
                             double* u_vals);
/* Number of dependency levels of the L and U sweeps. */
spk_status spk_ilu0_levels(spk_precond M, int64_t* levels_l,
                           int64_t* levels_u);
/* forward_solve(L, b) / backward_solve(U, b) with the factors of an ILU
 * preconditioner (triangular_solve.hpp:27-55, 67-98). */
spk_status spk_ilu0_forward(spk_precond M, int dtype, const void* d_b,
                            void* d_x);
spk_status spk_ilu0_backward(spk_precond M, int dtype, const void* d_b,
                             void* d_x);

/* --------------------------------------------------------------- solvers */

/* Mirrors CGConfig / Gmres_config / Fgmres_config / Bicgstab_config
 * (conjugate_gradient.hpp:47-74, gmres.hpp:24-47, fgmres.hpp:24-47,
 * bicgstab.hpp:24-42).  `restart` is restart_iterations for CG and
 * restart_dimension for (F)GMRES.  `precision` selects the arithmetic of
 * the Krylov vectors: SPK_F64 (reference) or SPK_F32.  b and x are device
 * vectors of that precision. */
typedef struct {
  double tolerance;
  int64_t restart;
  int64_t max_iterations;
  int32_t collect_residuals;
  int32_t precision;
} spk_solver_config;

/* Mirrors CGSummary / Gmres_summary / ... (conjugate_gradient.hpp:26-40).
 * iteration_residuals go to the caller's `history` buffer (host, at least
 * max_iterations doubles) when collect_residuals is set. */
typedef struct {
  double residual_norm;
  int64_t computed_iterations;
  int32_t converged;
  int32_t status; /* SPK_OK or SPK_BREAKDOWN */
  int64_t history_len;
} spk_summary;

/* conjugate_gradient(b, x, cfg, A, M_L, M_R) (conjugate_gradient.hpp:339-366).
 * NULL preconditioners are the identity. */
spk_status spk_cg_solve(spk_csr A, spk_precond M_left, spk_precond M_right,
                        const spk_solver_config* cfg, const void* d_b,
                        void* d_x, spk_summary* summary, double* history);
/* gmres(b, x, cfg, A, M_L, M_R) (gmres.hpp:400-427), modified Gram-Schmidt
 * and Givens rotations as the reference. */
spk_status spk_gmres_solve(spk_csr A, spk_precond M_left, spk_precond M_right,
                           const spk_solver_config* cfg, const void* d_b,
                           void* d_x, spk_summary* summary, double* history);
/* fgmres(b, x, cfg, A, M) (fgmres.hpp:452-470). */
spk_status spk_fgmres_solve(spk_csr A, spk_precond M,
                            const spk_solver_config* cfg, const void* d_b,
                            void* d_x, spk_summary* summary, double* history);
/* bicgstab(b, x, cfg, A, M_L, M_R) (bicgstab.hpp:383-410); counts matvecs. */
spk_status spk_bicgstab_solve(spk_csr A, spk_precond M_left,
                              spk_precond M_right,
                              const spk_solver_config* cfg, const void* d_b,
                              void* d_x, spk_summary* summary,
                              double* history);

/* Mixed-precision iterative refinement (no reference counterpart; the
 * closest structure is the outer correction loop of richardson.hpp:112):
 *   r = b - A x in FP64; stop when ||r|| / ||b|| < tolerance (FP64 true
 *   residual);  solve A d = r / ||r|| in FP32 with the inner Krylov method
 *   (FP32 vectors and SpMV, FP64 dot products, Arnoldi/Givens scalars) to
 *   inner_tolerance;  x += ||r|| d in FP64.
 * inner_method: SPK_IR_GMRES (M as right preconditioner), SPK_IR_FGMRES,
 * SPK_IR_CG (M as left preconditioner), SPK_IR_BICGSTAB (M right). */
enum { SPK_IR_GMRES = 0, SPK_IR_FGMRES = 1, SPK_IR_CG = 2, SPK_IR_BICGSTAB = 3 };

typedef struct {
  double tolerance;       /* FP64 true relative residual target */
  int64_t max_outer;      /* refinement steps */
  double inner_tolerance; /* relative tolerance of each FP32 inner solve */
  int64_t restart;        /* inner restart dimension / restart_iterations */
  int64_t max_inner;      /* inner iteration budget per refinement step */
  int32_t inner_method;
  int32_t pad;
} spk_ir_config;

typedef struct {
  double relres;                 /* final FP64 ||b - A x|| / ||b|| */
  int64_t outer_iterations;      /* refinement steps taken */
  int64_t inner_iterations;      /* total inner iterations */
  int32_t converged;
  int32_t status;
} spk_ir_summary;

/* history (host, max_outer + 1 doubles, may be NULL): the FP64 true
 * relative residual before each refinement step and at the end. */
spk_status spk_ir_solve(spk_csr A, spk_precond M, const spk_ir_config* cfg,
                        const double* d_b, double* d_x, spk_ir_summary* out,
                        double* history);

/* ||b - A x||_2 / ||b||_2 in FP64 (the true-residual check). */
spk_status spk_true_relres(spk_csr A, const double* d_b, const double* d_x,
                           double* relres);

/* ------------------------------------------------------------ generators */

/* splitmix64 uniform[-1,1] vector: element k = 2 u01(mix(seed + (k+1) g)) - 1
 * (the seeded right-hand sides of the benchmarks). */
spk_status spk_fill_uniform_pm1(uint64_t seed, int64_t n, double* d_out);
/* poisson_2d(nx, ny, shift) (matgen.hpp:173-265) built directly on the
 * device; shift = 0 is the singular Neumann Laplacian of the reference. */
spk_status spk_gen_poisson_2d(int64_t nx, int64_t ny, double shift,
                              spk_csr* out);
/* 7-point 3D Dirichlet Laplacian (diagonal 6, off-diagonal -1); no
 * reference counterpart. */
spk_status spk_gen_poisson_3d(int64_t nx, int64_t ny, int64_t nz,
                              spk_csr* out);
/* 7-point 3D upwind convection-diffusion extending convection_diffusion_2d
 * (matgen.hpp:274-325) to z with a Dirichlet diagonal 6d + cx + cy + cz. */
spk_status spk_gen_convdiff_3d(int64_t nx, int64_t ny, int64_t nz, double d,
                               double cx, double cy, double cz, spk_csr* out);
/* Power-law SPD matrix: row lengths from a Zipf(1.5) law on [3, 4096],
 * uniform columns, values uniform [-1, 0), symmetric part (A + A^T)/2 on the
 * union pattern, diagonal = sum |off-diagonal| + 1 (random_sparse's
 * dominance rule, matgen.hpp:657-658).  Built on the device (sort, unique,
 * scan); no reference counterpart. */
spk_status spk_gen_powerlaw_spd(int64_t n, uint64_t seed, spk_csr* out);

#ifdef __cplusplus
}
#endif

#endif /* SPARKIT_HIP_H */
