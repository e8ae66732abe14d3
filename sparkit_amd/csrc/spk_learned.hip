// spk_learned.hip -- the learned preconditioner's solve with a fallback:
// FGMRES(m) with an explicit approximate inverse M in its preconditioner
// slot (fgmres.hpp:452-470), guarded by a check of M, a stall monitor on
// the restart cycles and the FP64 true residual; on failure Jacobi-PCG
// (conjugate_gradient.hpp:122-312, jacobi.hpp:25-50) for symmetric A or
// ILU(0)-GMRES(m) (gmres.hpp:99-375, ilu.hpp:35-179) otherwise.
//
// The checks are two reductions over M's store (non-finite values; rows
// whose diagonal entry is missing, not finite or not positive) and one over
// A's (entries whose mirror a_ji is missing or differs: A is symmetric when
// there are none); the monitor reads the residual estimate the FGMRES host
// loop already downloads once per cycle.
#include <cmath>
#include <limits>

#include "spk_solver_common.hpp"

namespace spk {
namespace {

enum : int { kWsLearnedX0 = 122, kWsLearnedCnt = 123 };

// counts: [0] non-finite values of M, [1] rows with a bad diagonal
void check_inverse(spk_csr_s& M, int64_t* bad_values, int64_t* bad_diag) {
  const int64_t n = M.nrows;
  const int32_t* rp = M.rp.p;
  const int32_t* ci = M.ci.p;
  const double* v = M.v64.p;
  double* cnt = (double*)workspace().get(kWsLearnedCnt, 2 * sizeof(double));
  launch_reduce<2>(n, [=] __device__(int64_t i, double* acc) {
    const int a = rp[i], e = rp[i + 1];
    bool diag_ok = false;
    for (int k = a; k < e; ++k) {
      const double m = v[k];
      if (!isfinite(m)) acc[0] += 1.0;
      if (ci[k] == (int32_t)i) diag_ok = isfinite(m) && m > 0.0;
    }
    if (!diag_ok) acc[1] += 1.0;
  }, [=] __device__(const double* t) {
    cnt[0] = t[0];
    cnt[1] = t[1];
  });
  double h[2];
  SPK_HIP_TRY(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, stream()));
  SPK_HIP_TRY(hipStreamSynchronize(stream()));
  *bad_values = (int64_t)h[0];
  *bad_diag = (int64_t)h[1];
}

// entries a_ij whose mirror a_ji is missing or holds a different value
// (columns are sorted within a row: binary search)
int64_t asymmetric_entries(spk_csr_s& A) {
  const int64_t n = A.nrows;
  const int32_t* rp = A.rp.p;
  const int32_t* ci = A.ci.p;
  const double* v = A.v64.p;
  double* cnt = (double*)workspace().get(kWsLearnedCnt, 2 * sizeof(double));
  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = ci[k];
      int lo = rp[j], hi = rp[j + 1];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ci[mid] < (int32_t)i) lo = mid + 1;
        else hi = mid;
      }
      if (lo >= rp[j + 1] || ci[lo] != (int32_t)i || v[lo] != v[k]) acc[0] += 1.0;
    }
  }, [=] __device__(const double* t) { cnt[0] = t[0]; });
  double h = 0.0;
  SPK_HIP_TRY(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, stream()));
  SPK_HIP_TRY(hipStreamSynchronize(stream()));
  return (int64_t)h;
}

double true_relres(spk_csr_s* A, const double* b, const double* x) {
  double r = 0.0;
  const spk_status st = spk_true_relres(A, b, x, &r);
  SPK_REQUIRE(st == SPK_OK, st, std::string("true residual: ") + spk_last_error());
  return r;
}

void check_call(spk_status st, const char* what) {
  if (st != SPK_OK && st != SPK_BREAKDOWN)
    throw Error{st, std::string(what) + ": " + spk_last_error()};
}

// A preconditioner handle destroyed at scope exit.
struct PrecondGuard {
  spk_precond p = nullptr;
  ~PrecondGuard() {
    if (p) (void)spk_precond_destroy(p);
  }
};

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_learned_solve(spk_csr A, spk_precond M,
                                        const spk_learned_config* cfg,
                                        const double* d_b, double* d_x,
                                        spk_learned_summary* out) {
  return guarded([&] {
    TraceRange trace_("spk_learned_solve");
    SPK_REQUIRE(A && cfg && d_b && d_x && out, SPK_INVALID_ARGUMENT,
                "null argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "solver requires a square operator");
    SPK_REQUIRE(!M || (M->kind == spk_precond_s::kSpmv && M->n == A->nrows),
                SPK_INVALID_ARGUMENT,
                "M must be an explicit approximate inverse of A's size");
    SPK_REQUIRE(cfg->tolerance > 0.0 && cfg->restart > 0 &&
                    cfg->max_iterations >= 0 && cfg->fallback_max_iterations >= 0,
                SPK_INVALID_ARGUMENT, "bad tolerance, restart or budget");
    SPK_REQUIRE(cfg->stall_cycles >= 1 && cfg->stall_factor > 0.0 &&
                    cfg->stall_factor <= 1.0,
                SPK_INVALID_ARGUMENT, "stall_cycles >= 1, stall_factor in (0, 1]");
    SPK_REQUIRE(cfg->fallback >= SPK_FALLBACK_AUTO && cfg->fallback <= SPK_FALLBACK_NONE,
                SPK_INVALID_ARGUMENT, "unknown fallback");
    SPK_REQUIRE(cfg->orthogonalization == SPK_ORTHO_MGS ||
                    cfg->orthogonalization == SPK_ORTHO_CGS2 ||
                    cfg->orthogonalization == SPK_ORTHO_CGS1,
                SPK_INVALID_ARGUMENT, "unknown orthogonalization");
    *out = spk_learned_summary{};
    out->learned_relres = std::numeric_limits<double>::quiet_NaN();
    const int64_t n = A->nrows;
    const size_t vb = sizeof(double) * (size_t)(n > 0 ? n : 1);
    hipStream_t s = stream();
    double bb = 0.0;
    check_call(spk_dot(SPK_F64, n, d_b, d_b, &bb), "||b||");
    if (bb == 0.0) {  // b = 0: x = 0, as the reference's solvers return
      SPK_HIP_TRY(hipMemsetAsync(d_x, 0, vb, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      out->method = SPK_METHOD_LEARNED_FGMRES;
      out->converged = 1;
      out->relres = 0.0;
      return;
    }
    double* x0 = (double*)workspace().get(kWsLearnedX0, vb);
    SPK_HIP_TRY(hipMemcpyAsync(x0, d_x, vb, hipMemcpyDeviceToDevice, s));
    const double rel0 = true_relres(A, d_b, d_x);
    if (rel0 == 0.0) {  // x0 is exact
      out->method = SPK_METHOD_LEARNED_FGMRES;
      out->converged = 1;
      out->relres = 0.0;
      return;
    }

    // 1. check M, 2. FGMRES under the stall monitor, 3. the true residual
    if (!M) {
      out->reason |= SPK_REASON_NO_PRECONDITIONER;
    } else {
      {
        TraceRange check_("learned solve: check M");
        check_inverse(*M->M, &out->bad_values, &out->bad_diagonal);
      }
      if (out->bad_values) out->reason |= SPK_REASON_BAD_VALUES;
      if (out->bad_diagonal) out->reason |= SPK_REASON_BAD_DIAGONAL;
    }
    if (out->reason == 0) {
      spk_solver_config c{cfg->tolerance, cfg->restart, cfg->max_iterations,
                          0, SPK_F64, cfg->orthogonalization, 0};
      double prev = std::numeric_limits<double>::infinity();
      int slow = 0;
      bool stalled = false;
      const CycleMonitor monitor = [&](int64_t, double res) {
        if (!std::isfinite(res)) {
          stalled = true;
          return false;
        }
        slow = res > cfg->stall_factor * prev ? slow + 1 : 0;
        prev = res;
        if (slow >= cfg->stall_cycles) {
          stalled = true;
          return false;
        }
        return true;
      };
      spk_summary sum{};
      TraceRange solve_("learned solve: FGMRES");
      fgmres_monitored(A, M, &c, d_b, d_x, &sum, monitor);
      check_preconds(M, nullptr);
      out->learned_iterations = sum.computed_iterations;
      out->learned_relres = true_relres(A, d_b, d_x);
      if (std::isfinite(out->learned_relres) &&
          out->learned_relres < cfg->tolerance) {
        out->method = SPK_METHOD_LEARNED_FGMRES;
        out->converged = 1;
        out->relres = out->learned_relres;
        return;
      }
      out->reason |= stalled ? SPK_REASON_STALLED : SPK_REASON_NOT_CONVERGED;
    }
    if (cfg->fallback == SPK_FALLBACK_NONE) {
      out->method = SPK_METHOD_LEARNED_FGMRES;
      out->relres = true_relres(A, d_b, d_x);
      return;
    }

    TraceRange fallback_("learned solve: fallback");
    // the fallback starts from the better of x0 and the learned x
    if (!(std::isfinite(out->learned_relres) && out->learned_relres < rel0))
      SPK_HIP_TRY(hipMemcpyAsync(d_x, x0, vb, hipMemcpyDeviceToDevice, s));
    int kind = cfg->fallback;
    if (kind == SPK_FALLBACK_AUTO)
      kind = asymmetric_entries(*A) == 0 ? SPK_FALLBACK_JACOBI_PCG
                                         : SPK_FALLBACK_ILU_GMRES;
    PrecondGuard P;
    if (kind == SPK_FALLBACK_ILU_GMRES &&
        spk_precond_create_ilu0(A, 0, &P.p) != SPK_OK)
      P.p = nullptr;  // no usable ILU(0): Jacobi, else none, below
    if (!P.p && spk_precond_create_jacobi(A, &P.p) != SPK_OK) P.p = nullptr;
    out->method = kind == SPK_FALLBACK_JACOBI_PCG ? SPK_METHOD_JACOBI_PCG
                                                  : SPK_METHOD_ILU_GMRES;
    double rel = std::numeric_limits<double>::infinity();
    // a solver's own test reads its recursive (CG) or estimated residual:
    // a run that stops just above the tolerance in the true residual is
    // continued from its x (at most three runs)
    for (int round = 0; round < 3 && !(rel < cfg->tolerance); ++round) {
      const int64_t left = cfg->fallback_max_iterations - out->fallback_iterations;
      if (left <= 0) break;
      spk_summary sum{};
      if (kind == SPK_FALLBACK_JACOBI_PCG) {
        spk_solver_config c{cfg->tolerance, left, left, 0, SPK_F64, 0, 0};
        check_call(spk_cg_solve(A, P.p, nullptr, &c, d_b, d_x, &sum, nullptr),
                   "fallback Jacobi-PCG");
      } else {
        spk_solver_config c{cfg->tolerance, cfg->restart, left, 0, SPK_F64,
                            SPK_ORTHO_CGS2, 0};
        check_call(spk_gmres_solve(A, nullptr, P.p, &c, d_b, d_x, &sum, nullptr),
                   "fallback ILU(0)-GMRES");
      }
      out->fallback_iterations += sum.computed_iterations;
      rel = true_relres(A, d_b, d_x);
      if (sum.computed_iterations == 0) break;
    }
    out->relres = rel;
    out->converged = rel < cfg->tolerance ? 1 : 0;
  });
}
