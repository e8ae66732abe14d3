// spk_ir.hip -- mixed-precision iterative refinement around the device
// Krylov solvers:
//   r = b - A x (FP64 SpMV + FP64 norm: the true residual)
//   while ||r|| / ||b|| >= tol:
//     r32 = r / ||r||  (FP32, unit norm, so the inner solve is well scaled)
//     d32 = 0; inner Krylov solve of A d = r32 in FP32 (SpMV, vectors) with
//       FP64 dots and scalars, to inner_tolerance
//     x += ||r|| d   (FP64)
//     r = b - A x
// The FP64 work per refinement step is one SpMV, one norm and one update;
// everything inside the inner solve streams FP32 values (8 bytes per
// nonzero instead of 12).
#include "spk_solver_common.hpp"

namespace spk {
namespace {

struct IrState {
  double rnorm, bnorm;
};

struct EpiResidualNorm {  // r = b - A x, partial |r|^2
  static constexpr bool kReduce = true;
  const double* __restrict__ b;
  double* __restrict__ r;
  IrState* st;
  const int* done = nullptr;
  __device__ double row(int i, double s) const {
    double ri = b[i] - s;
    r[i] = ri;
    return ri * ri;
  }
  __device__ void finalize(double t) const { st->rnorm = sqrt(t); }
};

void ir_solve(spk_csr_s* A, spk_precond_s* M, const spk_ir_config* cfg,
              const double* b, double* x, spk_ir_summary* out,
              double* history) {
  const int64_t n = A->nrows;
  Workspace& ws = workspace();
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double* r = (double*)ws.get(kWsIr + 0, nn * 8);
  float* r32 = (float*)ws.get(kWsIr + 1, nn * 4);
  float* d32 = (float*)ws.get(kWsIr + 2, nn * 4);
  DevState<IrState> st(kWsIr + 3);
  IrState* S = st.d;
  hipStream_t s = stream();

  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += b[i] * b[i];
  }, [=] __device__(const double* t) { S->bnorm = sqrt(t[0]); });
  launch_spmv(*A, (const double*)x, EpiResidualNorm{b, r, S}, s);
  IrState h = st.download();
  *out = spk_ir_summary{};
  if (h.bnorm == 0.0) {
    launch_map(n, [=] __device__(int64_t i) { x[i] = 0.0; });
    out->converged = 1;
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }
  spk_solver_config inner{};
  inner.tolerance = cfg->inner_tolerance;
  inner.restart = cfg->restart;
  inner.max_iterations = cfg->max_inner;
  inner.collect_residuals = 0;
  inner.precision = SPK_F32;
  inner.orthogonalization = cfg->orthogonalization;
  for (int64_t k = 0;; ++k) {
    double rel = h.rnorm / h.bnorm;
    if (history) history[k] = rel;
    out->relres = rel;
    out->outer_iterations = k;
    if (rel < cfg->tolerance) {
      out->converged = 1;
      break;
    }
    if (k >= cfg->max_outer) break;
    const double scale = 1.0 / h.rnorm;
    launch_map(n, [=] __device__(int64_t i) {
      r32[i] = (float)(r[i] * scale);
      d32[i] = 0.0f;
    });
    spk_summary is{};
    spk_status st_in = SPK_OK;
    switch (cfg->inner_method) {
      case SPK_IR_FGMRES:
        st_in = spk_fgmres_solve(A, M, &inner, r32, d32, &is, nullptr);
        break;
      case SPK_IR_CG:
        st_in = spk_cg_solve(A, M, nullptr, &inner, r32, d32, &is, nullptr);
        break;
      case SPK_IR_BICGSTAB:
        st_in = spk_bicgstab_solve(A, nullptr, M, &inner, r32, d32, &is, nullptr);
        break;
      default:
        st_in = spk_gmres_solve(A, nullptr, M, &inner, r32, d32, &is, nullptr);
    }
    if (st_in != SPK_OK) throw Error{st_in, spk_last_error()};
    out->inner_iterations += is.computed_iterations;
    const double rn = h.rnorm;
    launch_map(n, [=] __device__(int64_t i) { x[i] += rn * (double)d32[i]; });
    launch_spmv(*A, (const double*)x, EpiResidualNorm{b, r, S}, s);
    h = st.download();
    if (is.status == SPK_BREAKDOWN && is.computed_iterations == 0) {
      out->status = SPK_BREAKDOWN;
      out->relres = h.rnorm / h.bnorm;
      break;
    }
  }
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_ir_solve(spk_csr A, spk_precond M,
                                   const spk_ir_config* cfg, const double* d_b,
                                   double* d_x, spk_ir_summary* out,
                                   double* history) {
  return guarded([&] {
    SPK_REQUIRE(A && cfg && d_b && d_x && out, SPK_INVALID_ARGUMENT,
                "null argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "solver requires a square operator");
    SPK_REQUIRE(!M || M->n == A->nrows, SPK_INVALID_ARGUMENT,
                "preconditioner size mismatch");
    SPK_REQUIRE(cfg->max_outer >= 0 && cfg->max_inner >= 0, SPK_INVALID_ARGUMENT,
                "negative iteration budget");
    ir_solve(A, M, cfg, d_b, d_x, out, history);
    check_preconds(M, nullptr);
  });
}
