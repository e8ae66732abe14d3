// spk_solver_common.hpp -- pieces shared by the device Krylov solvers:
// workspace slot map, argument checks and the chunked issue loop.
#pragma once

#include <functional>

#include "spk_precond.hpp"
#include "spk_vec.hpp"

namespace spk {

// Host-side monitor of a restarted (F)GMRES solve: called after every
// restart cycle that did not converge, with the iterations so far and the
// residual estimate at the cycle's end; returning false ends the solve.
using CycleMonitor = std::function<bool(int64_t iters, double resnorm)>;

// FP64 FGMRES(m) (spk_fgmres_solve) under a cycle monitor.
void fgmres_monitored(spk_csr_s* A, spk_precond_s* M,
                      const spk_solver_config* cfg, const double* b, double* x,
                      spk_summary* out, const CycleMonitor& monitor);

// Workspace slots (each solver owns a range so vectors stay allocated
// between calls and the timed region performs no hipMalloc).
enum : int {
  kWsCg = 0,         // 0..9
  kWsGmres = 10,     // 10..29
  kWsBicgstab = 30,  // 30..44
  kWsIr = 45,        // 45..59
  kWsMisc = 60,
  kWsCgs = 64,       // 64..74
  kWsBicg = 75,      // 75..84
  kWsMinres = 85,    // 85..94
  kWsQmr = 95,       // 95..109
  kWsCgsParts = 110,   // 110..112: CGS hand-off partials
  kWsBicgParts = 113,  // 113..115: BiCG hand-off partials
  kWsMinresParts = 116,  // 116..117: MINRES hand-off partials
  kWsQmrParts = 118,     // 118..121: QMR hand-off partials
};


inline bool is_identity(const spk_precond_s* M) {
  return M == nullptr || M->kind == spk_precond_s::kIdentity;
}
inline bool is_jacobi(const spk_precond_s* M) {
  return M != nullptr && M->kind == spk_precond_s::kJacobi;
}
// The diagonal of a Jacobi preconditioner, nullptr for the identity.
template <typename T>
inline const T* diag_of(spk_precond_s* M) {
  return is_jacobi(M) ? M->inv_diag<T>() : nullptr;
}
inline bool is_diagonal(const spk_precond_s* M) {
  return is_identity(M) || is_jacobi(M);
}

template <typename T>
inline void apply_precond(spk_precond_s* M, int64_t n, const T* in, T* out) {
  if (is_identity(M)) {
    if (in != out)
      SPK_HIP_TRY(hipMemcpyAsync(out, in, (size_t)n * sizeof(T),
                                 hipMemcpyDeviceToDevice, stream()));
  } else {
    M->apply(in, out);
  }
}

// Reports a triangular sweep whose bounded wait expired during the solve.
inline void check_preconds(spk_precond_s* a, spk_precond_s* b) {
  if (a) a->check_sweeps();
  if (b) b->check_sweeps();
}

inline void check_solver_args(spk_csr_s* A, spk_precond_s* ML,
                              spk_precond_s* MR, const spk_solver_config* cfg,
                              const void* b, void* x) {
  SPK_REQUIRE(A && cfg && b && x, SPK_INVALID_ARGUMENT, "null argument");
  SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
              "solver requires a square operator");
  SPK_REQUIRE(!ML || ML->n == A->nrows, SPK_INVALID_ARGUMENT,
              "left preconditioner size mismatch");
  SPK_REQUIRE(!MR || MR->n == A->nrows, SPK_INVALID_ARGUMENT,
              "right preconditioner size mismatch");
  SPK_REQUIRE(cfg->precision == SPK_F64 || cfg->precision == SPK_F32,
              SPK_INVALID_ARGUMENT, "precision must be SPK_F64 or SPK_F32");
  SPK_REQUIRE(cfg->max_iterations >= 0, SPK_INVALID_ARGUMENT,
              "max_iterations must be >= 0");
}

// Host issue loop: iterations are queued in chunks without reading anything
// back; after each chunk one small device-to-host copy tells whether the
// device-side convergence flag is set.  Iterations queued after convergence
// return immediately (their kernels test the flag).
template <typename IssueOne, typename Done>
void chunked_loop(int64_t max_it, IssueOne issue, Done is_done,
                  int64_t first_chunk = 4, int64_t max_chunk = 64) {
  int64_t k = 0, chunk = first_chunk;
  while (k < max_it) {
    int64_t kend = std::min(max_it, k + chunk);
    for (; k < kend; ++k) issue(k);
    if (is_done()) break;
    chunk = std::min(chunk * 2, max_chunk);
  }
}

}  // namespace spk
