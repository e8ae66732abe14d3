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
  SPK_NO_DEVICE = 6,
  SPK_RUNTIME_ERROR = 7 /* std::runtime_error (matrix_market.hpp) */
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

/* Gather-throughput ceiling (no reference counterpart; a measurement for
 * the roofline of gather-bound SpMVs): streams `count` (int32 index, FP64
 * value) pairs with nontemporal loads and multiplies each value by
 * x[index], indices uniform (splitmix64) over a window of `window` FP64
 * entries, 2048 workgroups of 512 threads; returns the best of `reps`
 * timed launches as gathers per second and streamed bytes (12 per pair)
 * per second.  Allocates and frees its buffers (count * 12 + window * 8
 * bytes); count a multiple of 4096, window in [1, 2^31). */
spk_status spk_gather_rate(int64_t window, int64_t count, int reps,
                           double* gathers_per_s, double* stream_bytes_per_s);

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
/* y = A x and *d_result (a device double) = <x, y> summed in FP64 in a fixed
 * order, in one SpMV with the sum in its epilogue: compute_step_size's
 * apply-then-dot of conjugate_gradient.hpp:232-239 (square A). */
spk_status spk_csr_spmv_dot(spk_csr A, int dtype, const void* d_x, void* d_y,
                            double* d_result);

/* Assembly on the device from `count` (row, col[, value]) entries held in
 * device memory (int64 indices, the reference's Index): the pattern of
 * Compressed_row_sparsity(Shape, indices)
 * (Compressed_row_sparsity_Impl.cpp:10-40: sorted, duplicates collapsed)
 * and, with d_vals, make_matrix's values (matgen.hpp:25-52: for a repeated
 * position the last entry in input order wins); d_vals NULL gives zero
 * values.  SPK_LOGIC_ERROR for rows or cols <= 1 (Shape.cpp:8-10) or a
 * negative index (Index.cpp:8); SPK_INVALID_ARGUMENT for an index outside
 * the shape. */
spk_status spk_csr_assemble(int64_t nrows, int64_t ncols, int64_t count,
                            const int64_t* d_rows, const int64_t* d_cols,
                            const double* d_vals, spk_csr* out);
/* A^T as a new store (the assembly of the swapped entries): its SpMV is
 * multiply_transpose(A, x) (sparse_blas.hpp:36-52) with each output summed
 * in the reference's order. */
spk_status spk_csr_transpose(spk_csr A, spk_csr* out);
/* reverse_cuthill_mckee(A.sparsity()) (reordering.cpp:152-203) on the
 * device: d_perm (n int64, device) receives perm[old] = new, for
 * spk_csr_permute(A, perm, SPK_PERM_BOTH) (dperm).  Neighbours are ordered
 * by degree with ties in column order (a stable sort, as the reference's
 * std::sort orders rows of up to 16 neighbours).  The pseudo-peripheral
 * start node (reordering.cpp:112-150) is exposed for tests.  Square
 * matrices only (SPK_INVALID_ARGUMENT). */
spk_status spk_csr_rcm(spk_csr A, int64_t* d_perm);
spk_status spk_csr_pseudo_peripheral_node(spk_csr A, int64_t* node);

/* Matrix Market files (matrix_market.hpp:43-163, matrix_market.cpp:28-133).
 * Reading follows read_matrix_market + to_compressed_row: banner checks
 * (coordinate format only, no complex field: SPK_RUNTIME_ERROR), '%'
 * comment lines, 1-based indices, pattern entries = 1, symmetric files
 * mirror their off-diagonal entries (skew-symmetric and hermitian are read
 * as written, as the reference does), a repeated position keeps the last
 * entry; the store is assembled on the device (spk_csr_assemble).
 * Writing follows write_matrix_market: a "coordinate real general" banner,
 * "rows cols nnz", then "row col value" per entry in storage order with
 * the value printed like std::ostream's default (%g, 6 significant
 * digits).  The _text variants work on memory: for writing, *len receives
 * the text's length and the text (NUL-terminated) is copied when buf has
 * room for len + 1 bytes. */
spk_status spk_csr_read_matrix_market(const char* path, spk_csr* out);
spk_status spk_csr_read_matrix_market_text(const char* text, size_t len,
                                           spk_csr* out);
spk_status spk_csr_write_matrix_market(spk_csr A, const char* path);
spk_status spk_csr_write_matrix_market_text(spk_csr A, char* buf, size_t cap,
                                            size_t* len);

/* (M + M^T) / 2 on M's own pattern (no reference counterpart: the SPD part
 * of a nonsymmetric approximate inverse, usable in conjugate_gradient's
 * preconditioner slot).  The pattern must be structurally symmetric
 * (SPK_INVALID_ARGUMENT otherwise); entry (i,j) = 0.5 (m_ij + m_ji), so the
 * result is exactly symmetric. */
spk_status spk_csr_symmetric_part(spk_csr M, spk_csr* out);

/* Permutations on the device, perm (device, int64, perm[old] = new) of
 * nrows (rows, both) or ncols (columns) entries: rperm (mode
 * SPK_PERM_ROWS), cperm (SPK_PERM_COLS: columns renamed, rows re-sorted) and
 * dperm (SPK_PERM_BOTH, P A P^T) of permutation.hpp:41-148, as a new store.
 * SPK_INVALID_ARGUMENT if perm is not a permutation
 * (is_valid_permutation, permutation.cpp:15-26) or for dperm of a
 * non-square matrix. */
#define SPK_PERM_ROWS 1
#define SPK_PERM_COLS 2
#define SPK_PERM_BOTH 3
spk_status spk_csr_permute(spk_csr A, const int64_t* d_perm, int mode,
                           spk_csr* out);

/* Layout of the store for SpMV: `panels` column panels (1: the plain
 * store).  Each panel is a store of its own holding the entries of an equal
 * share of the columns; an SpMV runs the panels in column order, carrying
 * every row's running sum from one to the next, so each launch gathers from
 * an L2-sized window of x (for matrices whose rows scatter over all columns;
 * rows summed by one lane keep the reference's storage-order sum).  Panels
 * other than the first and the last keep only their non-empty rows.  The
 * values are copied when the panels are made: call again after changing
 * them.  No reference counterpart (Compressed_row_matrix has one layout). */
spk_status spk_csr_set_panels(spk_csr A, int64_t panels);
/* the layout code: panels (>= 1), or -(slots * 65536 + window) for window
 * tiles */
spk_status spk_csr_get_panels(spk_csr A, int64_t* panels);
/* Window tiles (no reference counterpart), for matrices whose rows reach
 * across all columns: rows in groups of 1024 * slots (slots = 8 or 16 rows
 * per thread), and per window of `window` columns (256..8192, a multiple
 * of 64) each thread's entries of its rows stored together; the SpMV stages
 * each x window in LDS, so every gather is an LDS read.  A row's sum runs
 * in column order with separately rounded products and sums: the
 * reference's storage-order sum (sparse_blas.hpp:26-31) when columns
 * ascend within rows (required: SPK_INVALID_ARGUMENT otherwise).  Rows
 * longer than max(256, 8 * windows) entries run in a store of their own
 * by the row kernel, which splits a long row over lanes.  Replaces column
 * panels; the values are copied. */
spk_status spk_csr_set_window_tiles(spk_csr A, int64_t slots, int64_t window);
/* No reference counterpart: FP32 SpMVs of A (plain layout) read its FP32
 * values rounded to FP16 (round to nearest even), half the value bytes -- for a
 * preconditioner store such as an approximate inverse applied in an FP32
 * solve, whose FP64 true residual absorbs the rounding.  FP64 SpMVs are
 * unaffected.  The FP16 copy is made here; SPK_INVALID_ARGUMENT if a finite
 * value overflows FP16.  enable = 0 returns to the FP32 values. */
spk_status spk_csr_set_half_values(spk_csr A, int enable);
/* No reference counterpart: SpMVs of A's plain layout read each column
 * index as a 16-bit offset from the first row of its row block (2 bytes per
 * index instead of 4, in every precision; results unchanged bit for bit).
 * For banded stores in natural order, such as stencil matrices: every
 * column must lie within 32767 of its block's first row
 * (SPK_INVALID_ARGUMENT otherwise, and the store keeps its 32-bit reads).
 * Column panels and window tiles are unaffected.  enable = 0 returns to the
 * 32-bit indices. */
spk_status spk_csr_set_short_indices(spk_csr A, int enable);
/* a layout by its code (as spk_csr_get_panels reports it) */
spk_status spk_csr_set_layout(spk_csr A, int64_t code);
/* times one FP64 SpMV for 1, 4, 8, 12, 16, 24, 32 and 64 panels (those
 * leaving at least 32768 columns per panel) and, for matrices of at least
 * 131072 columns and 32 entries per row on average, three window-tile
 * shapes; keeps the fastest.  tried / tried_us (may be NULL, room for 11)
 * list the candidates' layout codes and times */
spk_status spk_csr_tune_layout(spk_csr A, int64_t* chosen, int64_t* tried,
                               float* tried_us, int64_t* ntried);

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
/* Schedule of the ILU(0) triangular sweeps (no reference counterpart; the
 * arithmetic, and so the result, is the same for both): SPK_ILU_AUTO (0,
 * the default: per-level launches for at least 65536 rows and at most 4096
 * levels, else sync-free), SPK_ILU_LEVELS (1: one launch per dependency
 * level), SPK_ILU_SYNC_FREE (2: one launch per sweep, rows wait for their
 * dependencies' published values). */
#define SPK_ILU_AUTO 0
#define SPK_ILU_LEVELS 1
#define SPK_ILU_SYNC_FREE 2
spk_status spk_ilu0_set_schedule(spk_precond M, int schedule);
spk_status spk_ilu0_get_schedule(spk_precond M, int* schedule);
/* Approximate ILU(0) apply (no reference counterpart; the reference's
 * ilu_apply, ilu.hpp:173-179, solves both triangles exactly, which stays the
 * default, sweeps = 0).  With sweeps = k > 0 each triangular solve becomes k
 * Jacobi iterations started from the diagonal-scaled right-hand side:
 *   y_0 = r,       y_j = r - L_s y_{j-1}
 *   z_0 = y_k / d, z_j = (y_k - U_s z_{j-1}) / d      (d: U's diagonal)
 * with L_s, U_s the strict triangles -- a fixed linear operator (k + 1
 * terms of each factor's Neumann series, equal to the exact solve once k
 * reaches the dependency depth), made of 2k + 1 bandwidth-bound passes
 * instead of two latency-bound sweeps.  spk_ilu0_forward/backward stay
 * exact.  k in [0, 1024]. */
spk_status spk_ilu0_set_jacobi_sweeps(spk_precond M, int sweeps);
/* No reference counterpart: the Jacobi-sweep apply's FP32 passes read the
 * strict factors' FP32 values rounded to FP16 (spk_csr_set_half_values on
 * L_s and U_s; U's diagonal stays FP32), for FP32 inner solves under an
 * FP64 refinement.  FP64 applies and the exact sweeps are unaffected. */
spk_status spk_ilu0_set_half_factors(spk_precond M, int enable);
spk_status spk_ilu0_get_half_factors(spk_precond M, int* enabled);
spk_status spk_ilu0_get_jacobi_sweeps(spk_precond M, int* sweeps);
/* Identity slot (conjugate_gradient_test.cpp:45-47). */
spk_status spk_precond_create_identity(int64_t n, spk_precond* out);
/* z = M r for an explicit sparse approximate inverse M (a store). */
spk_status spk_precond_create_spmv(spk_csr M, spk_precond* out);
/* A Krylov preconditioner for fgmres's flexible slot (fgmres.hpp:246-256;
 * no reference counterpart): z = the iterate after up to `iterations`
 * steps of conjugate_gradient (conjugate_gradient.hpp:122-312) on A z = r
 * from z = 0 with M_inner (NULL: identity; e.g. a learned approximate
 * inverse) in its preconditioner slot, stopping early once ||r - A z|| <
 * tolerance ||r||.  `precision` is the inner solve's (SPK_F32: FP32
 * vectors and SpMV, FP64 dot products and scalars); r and z are converted
 * at the boundary.  Not a fixed operator: conjugate_gradient rejects it.
 * A and M_inner must outlive it. */
spk_status spk_precond_create_inner_cg(spk_csr A, spk_precond M_inner,
                                       int64_t iterations, double tolerance,
                                       int precision, spk_precond* out);
/* Inner steps run by all applications of an inner-PCG preconditioner (or of
 * an inner mixed-precision solve, spk_precond_create_inner_ir) */
spk_status spk_precond_inner_cg_iterations(spk_precond M, int64_t* total);
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
 * preconditioner (triangular_solve.hpp:27-55, 67-98).  The reference
 * returns a new vector; here x is the caller's output and must not overlap
 * b (SPK_INVALID_ARGUMENT): the sync-free sweeps mark x before reading b. */
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
  /* (F)GMRES only: SPK_ORTHO_MGS (0, the reference's modified Gram-Schmidt,
   * gmres.hpp:258-267), SPK_ORTHO_CGS2 (classical Gram-Schmidt twice:
   * the same Krylov basis up to rounding, nine launches per Arnoldi step
   * instead of 2(j+2)+1) or SPK_ORTHO_CGS1 (no reference counterpart:
   * classical Gram-Schmidt once, half CGS2's passes over the basis; the
   * basis then keeps its orthogonality only as far as the problem's
   * conditioning allows, and the residual estimate the convergence test
   * reads can drift from the true residual -- for use where a true residual
   * decides, as under spk_ir_solve's FP64 refinement) */
  int32_t orthogonalization;
  int32_t reserved;
} spk_solver_config;
#define SPK_ORTHO_MGS 0
#define SPK_ORTHO_CGS2 1
#define SPK_ORTHO_CGS1 2

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
/* spk_cg_solve with no device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results as spk_cg_solve, but every work vector, the solver state and
 * the reduction partials live in one static 64 MiB device array of the
 * library (about 1.5M rows in FP64, 3M in FP32; larger solves, and
 * preconditioners other than identity, Jacobi or an explicit approximate
 * inverse, return SPK_INVALID_ARGUMENT before any work is queued).  What the
 * solve would otherwise make lazily -- a store's FP32 copy, a Jacobi
 * preconditioner's FP32 copy, a column-panel store's running sums -- must
 * already exist (one ordinary FP32 solve or SpMV makes them).  All calls on
 * one GPU share that GPU's array: calls that overlap on one GPU (other
 * threads, other streams) race and must be serialised by the caller. */
spk_status spk_cg_solve_noalloc(spk_csr A, spk_precond M_left,
                                spk_precond M_right,
                                const spk_solver_config* cfg, const void* d_b,
                                void* d_x, spk_summary* summary,
                                double* history);
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

/* cgs(b, x, cfg, A, M_L, M_R) (cgs.hpp:383-395); counts matvecs (two per
 * step); a zero <r^, q^> or rho stops with SPK_BREAKDOWN (undetected in the
 * reference).  cfg->restart is ignored. */
spk_status spk_cgs_solve(spk_csr A, spk_precond M_left, spk_precond M_right,
                         const spk_solver_config* cfg, const void* d_b,
                         void* d_x, spk_summary* summary, double* history);

/* bicg(b, x, cfg, A, A^T, M_L, M_R) (bicg.hpp:86-397): the transpose
 * operator is the store At (spk_csr_transpose(A)); counts matvecs (two per
 * step); zero <p~, q> or rho stop with SPK_BREAKDOWN. */
spk_status spk_bicg_solve(spk_csr A, spk_csr At, spk_precond M_left,
                          spk_precond M_right, const spk_solver_config* cfg,
                          const void* d_b, void* d_x, spk_summary* summary,
                          double* history);

/* minres(b, x, cfg, A, M_L, M_R) (minres.hpp:80-395): Lanczos with the
 * combined preconditioner M_R(M_L(.)), Givens QR, d-vector updates; the
 * residual estimate |phi_bar| is tested against ||b|| (:313-320); one
 * iteration per matvec.  rho == 0 in the QR update stops with
 * SPK_BREAKDOWN.  cfg->restart is ignored. */
spk_status spk_minres_solve(spk_csr A, spk_precond M_left, spk_precond M_right,
                            const spk_solver_config* cfg, const void* d_b,
                            void* d_x, spk_summary* summary, double* history);

/* qmr(b, x, cfg, A, A^T, M_L, M_R) (qmr.hpp:86-491): QMR without
 * look-ahead, the transpose operator as the store At; counts matvecs (two
 * per step, :251); the residual estimate is |tau|.  delta == 0 (the
 * reference repeats the step forever) and a zero Lanczos norm without
 * convergence stop with SPK_BREAKDOWN.  cfg->restart is ignored. */
spk_status spk_qmr_solve(spk_csr A, spk_csr At, spk_precond M_left,
                         spk_precond M_right, const spk_solver_config* cfg,
                         const void* d_b, void* d_x, spk_summary* summary,
                         double* history);

/* Mixed-precision iterative refinement (no reference counterpart; the
 * closest structure is the outer correction loop of richardson.hpp:112):
 *   r = b - A x in FP64; stop when ||r|| / ||b|| < tolerance (FP64 true
 *   residual);  solve A d = r / ||r|| in FP32 with the inner Krylov method
 *   (FP32 vectors and SpMV, FP64 dot products, Arnoldi/Givens scalars) to
 *   inner_tolerance;  x += ||r|| d in FP64.
 * inner_method: SPK_IR_GMRES (M as right preconditioner), SPK_IR_FGMRES,
 * SPK_IR_CG (M as left preconditioner), SPK_IR_BICGSTAB (M right), or
 * SPK_IR_CG_RELIABLE: one FP32 preconditioned CG recurrence (FP64 dots)
 * that is not restarted per step; instead, whenever its residual has
 * dropped by inner_tolerance since the last update (or below tolerance),
 * x += the FP32 correction in FP64, the residual is replaced by the FP64
 * true residual and the search direction is kept (reliable updates, van
 * der Vorst and Ye).  For it max_outer bounds the updates, max_inner the
 * total FP32 steps, and restart is unused; a Jacobi, identity or explicit
 * approximate-inverse (spk_precond_create_spmv) M keeps every kernel of a
 * step device-gated, other preconditioners are applied ungated. */
enum { SPK_IR_GMRES = 0, SPK_IR_FGMRES = 1, SPK_IR_CG = 2, SPK_IR_BICGSTAB = 3,
       SPK_IR_CG_RELIABLE = 4 };

typedef struct {
  double tolerance;       /* FP64 true relative residual target */
  int64_t max_outer;      /* refinement steps */
  double inner_tolerance; /* relative tolerance of each FP32 inner solve */
  int64_t restart;        /* inner restart dimension / restart_iterations */
  int64_t max_inner;      /* inner iteration budget per refinement step */
  int32_t inner_method;
  int32_t orthogonalization; /* inner (F)GMRES: SPK_ORTHO_MGS / _CGS2 / _CGS1 */
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

/* A preconditioner for fgmres's flexible slot (fgmres.hpp:246-256; no
 * reference counterpart) that runs a mixed-precision solve: z = the result
 * of spk_ir_solve(A, M_inner, cfg) on A z = r from z = 0, i.e. z reaches
 * ||r - A z|| < cfg->tolerance ||r|| (FP64 true residual) or cfg's budgets.
 * inner_method is SPK_IR_CG, SPK_IR_CG_RELIABLE or SPK_IR_BICGSTAB (an
 * inner (F)GMRES would share the outer one's work vectors); r and z are
 * FP64 at the boundary (converted for an FP32 caller).  Not a fixed
 * operator: conjugate_gradient and spk_ir_solve reject it.  A and M_inner
 * must outlive it; spk_precond_inner_cg_iterations counts its inner steps. */
spk_status spk_precond_create_inner_ir(spk_csr A, spk_precond M_inner,
                                       const spk_ir_config* cfg,
                                       spk_precond* out);

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

/* 2D anisotropic variable-coefficient diffusion on the unit square, bilinear
 * (Q1) finite elements, n_side x n_side interior nodes, Dirichlet boundary
 * eliminated: a 9-point SPD matrix with (3 n_side - 2)^2 nonzeros.  Per-cell
 * conductivity lognormal(0, sigma); per-matrix anisotropy eps log-uniform in
 * [1, 1e3] and rotation theta uniform in [0, pi) drawn from the seed
 * (isotropic != 0 forces eps = 1, theta = 0; sigma = 0 gives unit cells).
 * The learned-preconditioner training and test family; no reference
 * counterpart (the reference's 2D generators are the 5-point finite
 * difference stencils of matgen.hpp:173-386). */
spk_status spk_gen_aniso_q1_2d(int64_t n_side, uint64_t seed, double sigma,
                               int isotropic, spk_csr* out);

/* ------------------------------------------------------ learned SPAI
 * A small MLP (24 per-row features -> 64 -> 64 -> 9, ReLU) predicts the rows
 * of a sparse approximate inverse M on A's own pattern; M is applied as one
 * SpMV (spk_precond_create_spmv), e.g. inside FGMRES.  The network is
 * trained on the device (bf16 MFMA forward/backward, fp32 accumulation,
 * Adam on fp32 master weights) to minimise (1/n) ||I - M A||_F^2.  Rows must
 * hold a positive diagonal and at most 4 entries on each side of it
 * (SPK_INVALID_ARGUMENT otherwise).  No reference counterpart: the
 * reference's approximate inverses (neumann_preconditioner.hpp,
 * least_squares_preconditioner.hpp) are polynomial; the learned one takes
 * the same preconditioner slot (fgmres.hpp:452-470).
 * Parameters, canonical order: W1[24][64] b1[64] W2[64][64] b2[64]
 * W3[64][9] b3[9] (weights [in][out]).  A new model has He-uniform hidden
 * layers and a zero output layer, i.e. it predicts the Jacobi inverse. */
#define SPK_SPAI_NPARAM 6345
typedef struct spk_spai_model_s* spk_spai_model;
typedef struct {
  int64_t epochs;          /* passes over the training matrices */
  double learning_rate;    /* Adam step size */
  int64_t batch_matrices;  /* matrices per Adam step */
  uint64_t shuffle_seed;   /* per-epoch matrix order */
} spk_spai_train_config;

spk_status spk_spai_model_create(uint64_t seed, spk_spai_model* out);
spk_status spk_spai_model_destroy(spk_spai_model m);
spk_status spk_spai_model_get_params(spk_spai_model m, float* params);
/* replaces the weights and resets the optimiser state */
spk_status spk_spai_model_set_params(spk_spai_model m, const float* params);
/* No reference counterpart (the reference has no checkpoint; its only
 * resume is the warm-start x): the training state -- FP32 weights, Adam's
 * two moment vectors and step count -- to a checkpoint file and back, so a
 * loaded model predicts the same M and continues training exactly as the
 * saved one would.  Binary, little-endian: "SPKSPAI" NUL, uint32 version
 * (1), uint32 SPK_SPAI_NPARAM, int64 step, 3 x SPK_SPAI_NPARAM floats
 * (weights, first and second moments, canonical order), 64-bit FNV-1a of
 * the preceding bytes.  File errors, a wrong magic, version, size or hash:
 * SPK_RUNTIME_ERROR, and the model is left as it was. */
spk_status spk_spai_model_save(spk_spai_model m, const char* path);
spk_status spk_spai_model_load(spk_spai_model m, const char* path);
/* The training objective of spk_spai_train and spk_spai_loss_grad (no
 * reference counterpart): SPK_SPAI_LOSS_INVERSE (the default)
 * (1/n) ||I - M A||_F^2, or SPK_SPAI_LOSS_SYMMETRIC
 * (1/n) ||I - S A||_F^2 with S = (M + M^T) / 2 -- the operator CG applies
 * (spk_csr_symmetric_part), for structurally symmetric patterns
 * (SPK_INVALID_ARGUMENT otherwise).  Predictions are M either way.  Not
 * part of a checkpoint. */
enum { SPK_SPAI_LOSS_INVERSE = 0, SPK_SPAI_LOSS_SYMMETRIC = 1 };
spk_status spk_spai_model_set_objective(spk_spai_model m, int objective);
spk_status spk_spai_model_get_objective(spk_spai_model m, int* objective);
/* the 24 features of every row into a device array (n x 24, float) */
spk_status spk_spai_features(spk_csr A, float* d_features);
/* loss (1/n)||I - M A||_F^2 and its gradient (host array of
 * SPK_SPAI_NPARAM floats, may be NULL) for one matrix; no update */
spk_status spk_spai_loss_grad(spk_spai_model m, spk_csr A, double* loss,
                              float* grad);
/* epochs x ceil(count / batch) Adam steps; loss_history (host, may be NULL)
 * receives each step's mean loss over its batch */
spk_status spk_spai_train(spk_spai_model m, const spk_csr* mats, int64_t count,
                          const spk_spai_train_config* cfg,
                          double* loss_history);
/* M on A's pattern as a new store (the caller destroys it) */
spk_status spk_spai_predict(spk_spai_model m, spk_csr A, spk_csr* M_out);

/* ------------------------------------ learned solve with a fallback
 * An explicit approximate inverse M (spk_precond_create_spmv, e.g. over
 * spk_spai_predict's store) in FGMRES(m)'s preconditioner slot
 * (fgmres.hpp:452-470; the slot is applied at fgmres.hpp:246-256), guarded,
 * with the FP64 true residual as the judge:
 *  1. M is checked before use: every value finite, every row's diagonal
 *     entry present, finite and positive;
 *  2. FGMRES runs in FP64 and is stopped early when `stall_cycles`
 *     restart cycles in a row each fail to bring the residual estimate
 *     below stall_factor times the previous cycle's, or when it turns
 *     non-finite;
 *  3. x is accepted when ||b - A x|| / ||b|| < tolerance (FP64).
 * Otherwise the fallback solves from x0 (or from the learned x when that
 * is finite and better than x0): Jacobi-PCG (conjugate_gradient.hpp:
 * 122-312 with jacobi.hpp:25-50, restart_iterations = max_iterations) when A
 * is symmetric, else GMRES(m) with ILU(0) in the right slot (gmres.hpp:
 * 99-375, ilu.hpp:35-179) -- or the one `fallback` names.  M may be NULL
 * (the fallback runs at once).  No reference counterpart: the reference has
 * no learned preconditioner and does not detect a failing one. */
enum {
  SPK_FALLBACK_AUTO = 0,       /* Jacobi-PCG if A is symmetric, else ILU-GMRES */
  SPK_FALLBACK_JACOBI_PCG = 1,
  SPK_FALLBACK_ILU_GMRES = 2,
  SPK_FALLBACK_NONE = 3        /* report the learned result only */
};
/* spk_learned_summary.method: the solver whose x is returned */
enum {
  SPK_METHOD_LEARNED_FGMRES = 0,
  SPK_METHOD_JACOBI_PCG = 1,
  SPK_METHOD_ILU_GMRES = 2
};
/* spk_learned_summary.reason: why the fallback ran (bits) */
enum {
  SPK_REASON_NO_PRECONDITIONER = 1, /* M was NULL */
  SPK_REASON_BAD_VALUES = 2,        /* M holds NaN or infinite values */
  SPK_REASON_BAD_DIAGONAL = 4,      /* a diagonal entry missing, <= 0 or not finite */
  SPK_REASON_STALLED = 8,           /* FGMRES stopped making progress */
  SPK_REASON_NOT_CONVERGED = 16     /* budget spent above the tolerance */
};
typedef struct {
  double tolerance;                /* FP64 true relative residual target */
  int64_t restart;                 /* FGMRES / fallback GMRES restart dimension */
  int64_t max_iterations;          /* FGMRES iterations */
  int64_t fallback_max_iterations; /* fallback solver iterations */
  double stall_factor;             /* in (0, 1] */
  int32_t stall_cycles;            /* >= 1 */
  int32_t fallback;                /* SPK_FALLBACK_* */
  int32_t orthogonalization;       /* FGMRES: SPK_ORTHO_MGS, _CGS2 or _CGS1 */
  int32_t pad;
} spk_learned_config;
typedef struct {
  int32_t method;              /* SPK_METHOD_* */
  int32_t reason;              /* SPK_REASON_* bits, 0 if none */
  int32_t converged;           /* relres < tolerance */
  int32_t pad;
  int64_t learned_iterations;  /* FGMRES iterations run */
  int64_t fallback_iterations; /* fallback iterations run */
  int64_t bad_values;          /* non-finite entries of M */
  int64_t bad_diagonal;        /* rows of M with a bad diagonal */
  double learned_relres;       /* FP64 true residual after FGMRES (NaN if not run) */
  double relres;               /* FP64 true relative residual of the returned x */
} spk_learned_summary;

/* b, x: FP64 device vectors; x holds x0 on entry and the solution on exit */
spk_status spk_learned_solve(spk_csr A, spk_precond M,
                             const spk_learned_config* cfg, const double* d_b,
                             double* d_x, spk_learned_summary* summary);

/* ------------------------------------------------------- batched PCG
 * Thousands of independent SPD systems, stored as the diagonal blocks of
 * one CSR store, solved by Jacobi-PCG in one persistent launch: each
 * workgroup takes the next system (largest first) from a device queue and
 * runs Jacobi-PCG in `precision` (SPK_F32: FP32 SpMV and vectors, FP64 dot
 * products) with reliable updates: whenever the recurrence residual has
 * dropped by inner_tolerance since the last update, the FP32 correction is
 * added to x in FP64 and the FP64 true residual b - A x replaces the
 * recurrence residual (the search direction is kept).  A system is done
 * when its FP64 true relative residual is below tolerance.  SPK_F64 is
 * plain FP64 Jacobi-PCG per system (conjugate_gradient.hpp:122-312 with
 * jacobi.hpp:25-50) judged on the true residual.  No reference
 * counterpart for the batching itself. */
typedef struct spk_batch_s* spk_batch;
typedef struct {
  double tolerance;        /* FP64 true relative residual per system */
  double inner_tolerance;  /* residual reduction between FP64 updates */
  int64_t max_outer;       /* FP64 true-residual updates per system */
  int64_t max_inner;       /* PCG iterations per system */
  int32_t precision;       /* SPK_F32 or SPK_F64 inner arithmetic */
  int32_t pad;
} spk_batch_config;
typedef struct {
  int32_t block;      /* threads per workgroup: 256, 512, 1024 */
  int32_t wg_per_cu;  /* persistent workgroups per CU */
  int32_t lds_rows;   /* systems up to this many rows keep p in LDS (capped
                         to 128 KB per CU at the solve's precision) */
  int32_t pad;
} spk_batch_tuning;
typedef struct {
  int64_t total_inner_iterations, max_inner_iterations;
  int64_t total_outer_iterations, converged;
  double max_relres;
} spk_batch_summary;

/* count systems of 2D 5-point variable-coefficient diffusion (cell-centred
 * finite volumes, Dirichlet, lognormal(0,1) conductivity per cell,
 * harmonic-mean faces) on side x side grids, side uniform in
 * [side_min, side_max], all from `seed`. */
spk_status spk_batch_create_diffusion(int64_t count, uint64_t seed,
                                      int64_t side_min, int64_t side_max,
                                      spk_batch* out);
/* a batch over a copy of a block-diagonal store; row_offsets has count + 1
 * entries; SPK_INVALID_ARGUMENT if an entry couples two systems or a
 * system holds 2^29 rows or more */
spk_status spk_batch_create(spk_csr A, int64_t count, const int64_t* row_offsets,
                            spk_batch* out);
spk_status spk_batch_destroy(spk_batch B);
/* matrix: the batch's own store (valid while the batch lives) */
spk_status spk_batch_info(spk_batch B, int64_t* count, int64_t* total_rows,
                          int64_t* nnz, spk_csr* matrix);
spk_status spk_batch_row_offsets(spk_batch B, int64_t* row_offsets);
spk_status spk_batch_set_tuning(spk_batch B, const spk_batch_tuning* t);
spk_status spk_batch_get_tuning(spk_batch B, spk_batch_tuning* t);
/* times every candidate launch shape (256/512/1024 threads x 1/2/4
 * workgroups per CU x LDS bin threshold 0/16384 rows, those that fit) on
 * `sample` systems spread over the size range (x is overwritten), keeps
 * the fastest; tried / tried_ms (may be NULL, room for 18) list them */
spk_status spk_batch_autotune(spk_batch B, const spk_batch_config* cfg,
                              const double* d_b, double* d_x, int64_t sample,
                              spk_batch_tuning* chosen, spk_batch_tuning* tried,
                              float* tried_ms, int64_t* ntried);
/* x (FP64, all systems side by side) from b; x is reset to zero first.
 * use_graph: capture the launch into a hipGraph on first use (per b, x,
 * configuration and tuning) and replay it afterwards.  out may be NULL
 * (then the call does not synchronise). */
spk_status spk_batch_pcg_solve(spk_batch B, const spk_batch_config* cfg,
                               const double* d_b, double* d_x, int use_graph,
                               spk_batch_summary* out);
/* per-system results of the last solve (host arrays of count entries, any
 * may be NULL) */
spk_status spk_batch_results(spk_batch B, int32_t* inner_iterations,
                             int32_t* outer_iterations, double* relres,
                             int32_t* converged);

#ifdef __cplusplus
}
#endif

#endif /* SPARKIT_HIP_H */
