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

// ---------------------------------------------------------------------
// SPK_IR_CG_RELIABLE: one FP32 preconditioned CG recurrence whose residual
// is replaced by the FP64 true residual at "reliable updates" (van der
// Vorst and Ye's scheme), instead of a fresh inner CG per refinement step.
//   x64: FP64 solution; xi: FP32 correction since the last update
//   FP32 step: q = A p; alpha = rz / <p,q>; xi += alpha p; r -= alpha q;
//              z = M r; beta = <r,z> / rz; p = z + beta p
//   update when |r| <= delta |r_last update| (delta = inner_tolerance) or
//   |r| / |b| < tolerance:  x64 += xi; xi = 0; r64 = b - A x64 (FP64);
//   r = (float) r64; z = M r; beta = <r,z> / rz; p = z + beta p -- the
//   search direction is kept, so the Krylov progress survives the update.
// Convergence is decided on the FP64 true residual at updates only.  The
// host queues FP32 steps in chunks; a device flag (hold) turns the queued
// steps into no-ops once an update is due, and the host then issues it.

struct RcgState {
  double bnorm, rnorm, rupd, rz, alpha, beta, tol, delta;
  int64_t iters, updates;
  int32_t hold, converged, pad0, pad1;
  // one-kernel hand-off: rz by step parity, and the step whose second
  // kernel ran (gates the third)
  double rzb[2];
  int64_t k2_step;
};

struct RcgGateRun {
  const int* hold;
  __device__ bool operator()() const { return !*hold; }
};
struct RcgGateStep {
  const RcgState* st;
  int64_t step;
  __device__ bool operator()() const { return st->k2_step == step; }
};

struct EpiRcgDot {  // q = A p, partial <p, q>
  static constexpr bool kReduce = true;
  const float* __restrict__ p;
  float* __restrict__ q;
  RcgState* st;
  const int* done;
  __device__ double row(int i, float s) const {
    q[i] = s;
    return (double)p[i] * (double)s;
  }
  __device__ void finalize(double pq) const { st->alpha = st->rz / pq; }
};

struct EpiRcgPrec {  // z = M r (an explicit approximate inverse), <r, z>
  static constexpr bool kReduce = true;
  const float* __restrict__ r;
  float* __restrict__ z;
  RcgState* st;
  const int* done;
  __device__ double row(int i, float s) const {
    z[i] = s;
    return (double)r[i] * (double)s;
  }
  int par = 0;  // parity slot for the hand-off's rz
  __device__ void finalize(double rz) const {
    st->beta = rz / st->rz;
    st->rz = rz;
    st->rzb[par] = rz;
  }
};

struct EpiRcgTrue {  // r64 = b - A x64, |r64|^2
  static constexpr bool kReduce = true;
  const double* __restrict__ b;
  double* __restrict__ r;
  RcgState* st;
  const int* done = nullptr;
  __device__ double row(int i, double s) const {
    double ri = b[i] - s;
    r[i] = ri;
    return ri * ri;
  }
  __device__ void finalize(double t) const { st->rnorm = sqrt(t); }
};

// after the FP32 residual norm of a step: is an update due?
__device__ inline bool rcg_due(RcgState* st, double rr) {
  const double rn = sqrt(rr);
  st->iters += 1;
  if (rn <= st->delta * st->rupd || rn / st->bnorm < st->tol) {
    st->hold = 1;
    return true;
  }
  return false;
}

void cg_reliable(spk_csr_s* A, spk_precond_s* M, const spk_ir_config* cfg,
                 const double* b, double* x, double* r, spk_ir_summary* out,
                 double* history, double bnorm, double rnorm) {
  const int64_t n = A->nrows;
  Workspace& ws = workspace();
  const size_t nn = (size_t)(n > 0 ? n : 1);
  float* r32 = (float*)ws.get(kWsIr + 1, nn * 4);
  float* xi = (float*)ws.get(kWsIr + 2, nn * 4);
  float* z = (float*)ws.get(kWsIr + 4, nn * 4);
  float* p = (float*)ws.get(kWsIr + 5, nn * 4);
  float* q = (float*)ws.get(kWsIr + 6, nn * 4);
  DevState<RcgState> st(kWsIr + 7);
  RcgState h{};
  h.bnorm = bnorm;
  h.rnorm = rnorm;
  h.rupd = rnorm;
  h.tol = cfg->tolerance;
  h.delta = cfg->inner_tolerance;
  h.rz = 1.0;  // the first "beta" below multiplies p = 0
  st.upload(h);
  RcgState* S = st.d;
  const int* hold = &S->hold;
  hipStream_t s = stream();
  const bool jac = is_jacobi(M), ident = is_identity(M);
  const bool spai = M && M->kind == spk_precond_s::kSpmv;
  const float* invd = jac ? M->inv_diag<float>() : nullptr;
  // the one-kernel hand-off (every block of the next kernel sums the
  // previous partials) when the partials are few
  const bool handoff = (jac || ident || spai) &&
                       spmv_partials_count(*A) <= kProMaxPartials &&
                       pro_grid(n) <= kProMaxPartials &&
                       (!spai || spmv_partials_count(*M->M) <= kProMaxPartials);
  double* part1 = nullptr;
  double* part2 = nullptr;
  double* part3 = nullptr;
  if (handoff) {
    const size_t pb = sizeof(double) * 2 * (size_t)kProMaxPartials;
    part1 = (double*)ws.get(kWsIr + 8, pb);
    part2 = (double*)ws.get(kWsIr + 9, pb);
    part3 = (double*)ws.get(kWsIr + 10, pb);
  }
  int64_t issued = 0;

  // r = (float) r64, xi = 0, z = M r, beta = <r,z> / rz, rz = <r,z>; then
  // p = z + beta p (not when the update converged)
  auto restart = [&](bool first) {
    const int par = (int)(issued & 1);
    if (first) launch_map(n, [=] __device__(int64_t i) { p[i] = 0.0f; });
    if (jac || ident) {
      launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
        const float ri = (float)r[i];
        const float zi = invd ? invd[i] * ri : ri;
        r32[i] = ri;
        z[i] = zi;
        xi[i] = 0.0f;
        acc[0] += (double)ri * (double)zi;
      }, [=] __device__(const double* t) {
        S->beta = t[0] / S->rz;
        S->rz = t[0];
        S->rzb[par] = t[0];
      });
    } else {
      launch_map(n, [=] __device__(int64_t i) {
        r32[i] = (float)r[i];
        xi[i] = 0.0f;
      });
      if (spai) {
        launch_spmv(*M->M, (const float*)r32, EpiRcgPrec{r32, z, S, nullptr, par}, s);
      } else {
        M->apply((const float*)r32, z);
        launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
          acc[0] += (double)r32[i] * (double)z[i];
        }, [=] __device__(const double* t) {
          S->beta = t[0] / S->rz;
          S->rz = t[0];
          S->rzb[par] = t[0];
        });
      }
    }
    launch_map(n, [=] __device__(int64_t i) {
      p[i] = z[i] + (float)S->beta * p[i];
    }, hold);
  };

  // One step through the hand-off: K1 q = A p (partials <p,q>); K2 sums
  // them itself -> alpha, r -= alpha q, partials |r|^2 (and <r,z> for a
  // diagonal M); [K3 z = M r, partials <r,z>]; the last pass sums the
  // partials -> update due?, beta; xi += alpha p; p = z + beta p unless an
  // update is due (then the restart after it forms p).  Block 0 records
  // the scalars; rz of step k lives in rzb[k & 1].
  auto step_handoff = [&]() {
    const int par = (int)(issued & 1);
    const int64_t sid = issued + 1;
    const int c1 = launch_spmv_partials(*A, (const float*)p, EpiRcgDot{p, q, S, hold},
                                        part1, s);
    auto pro_alpha = [=] __device__(const double* t, double* sc, bool b0) {
      const double alpha = S->rzb[par] / t[0];
      sc[0] = alpha;
      if (b0) {
        S->alpha = alpha;
        S->k2_step = sid;
      }
    };
    // rcg_due's test and the new beta, once per block; block 0 records
    // them (nothing this kernel reads)
    auto pro_beta = [=] __device__(const double* t, double* sc, bool b0) {
      const double rn = sqrt(t[0]);
      const bool due = rn <= S->delta * S->rupd || rn / S->bnorm < S->tol;
      const double beta = t[1] / S->rzb[par];
      sc[0] = S->alpha;
      sc[1] = beta;
      sc[2] = due ? 1.0 : 0.0;
      if (b0) {
        S->iters += 1;
        if (due) {
          S->hold = 1;
          S->rz = S->rzb[par];
        } else {
          S->beta = beta;
          S->rz = t[1];
          S->rzb[par ^ 1] = t[1];
        }
      }
    };
    if (!spai) {
      const int g2 = launch_reduce_pro<1, 0, 2>(
          n, part1, c1, (const double*)nullptr, 0, RcgGateRun{hold}, pro_alpha,
          [=] __device__(int64_t i) {
            return Pack4<float>{r32[i], q[i], invd ? invd[i] : 1.0f, 0.0f};
          },
          [=] __device__(int64_t i, const Pack4<float>& e, double* acc, const double* sc) {
            const float a = (float)sc[0];
            const float ri = e.a - a * e.b;
            r32[i] = ri;
            acc[0] += (double)ri * (double)ri;
            acc[1] += (double)ri * (double)(invd ? e.c * ri : ri);
          }, part2);
      launch_reduce_pro<2, 0, 0>(
          n, part2, g2, (const double*)nullptr, 0, RcgGateStep{S, sid}, pro_beta,
          [=] __device__(int64_t i) {
            return Pack4<float>{p[i], xi[i], r32[i], invd ? invd[i] : 1.0f};
          },
          [=] __device__(int64_t i, const Pack4<float>& e, double*, const double* sc) {
            xi[i] = (float)sc[0] * e.a + e.b;
            if (sc[2] == 0.0) p[i] = (invd ? e.d * e.c : e.c) + (float)sc[1] * e.a;
          }, (double*)nullptr);
    } else {
      const int g2 = launch_reduce_pro<1, 0, 1>(
          n, part1, c1, (const double*)nullptr, 0, RcgGateRun{hold}, pro_alpha,
          [=] __device__(int64_t i) { return Pack4<float>{r32[i], q[i], 0.0f, 0.0f}; },
          [=] __device__(int64_t i, const Pack4<float>& e, double* acc, const double* sc) {
            const float a = (float)sc[0];
            const float ri = e.a - a * e.b;
            r32[i] = ri;
            acc[0] += (double)ri * (double)ri;
          }, part2);
      const int c3 = launch_spmv_partials(*M->M, (const float*)r32,
                                          EpiRcgPrec{r32, z, S, hold}, part3, s);
      launch_reduce_pro<1, 1, 0>(
          n, part2, g2, part3, c3, RcgGateStep{S, sid}, pro_beta,
          [=] __device__(int64_t i) { return Pack4<float>{p[i], xi[i], z[i], 0.0f}; },
          [=] __device__(int64_t i, const Pack4<float>& e, double*, const double* sc) {
            xi[i] = (float)sc[0] * e.a + e.b;
            if (sc[2] == 0.0) p[i] = e.c + (float)sc[1] * e.a;
          }, (double*)nullptr);
    }
  };

  auto step = [&]() {
    if (handoff) {
      step_handoff();
      return;
    }
    launch_spmv(*A, (const float*)p, EpiRcgDot{p, q, S, hold}, s);
    if (jac || ident) {
      launch_reduce<2>(n, [=] __device__(int64_t i, double* acc) {
        const float a = (float)S->alpha;
        xi[i] = a * p[i] + xi[i];
        const float ri = r32[i] - a * q[i];
        r32[i] = ri;
        acc[0] += (double)ri * (double)ri;
        acc[1] += (double)ri * (double)(invd ? invd[i] * ri : ri);
      }, [=] __device__(const double* t) {
        if (!rcg_due(S, t[0])) {
          S->beta = t[1] / S->rz;
          S->rz = t[1];
        }
      }, hold);
      launch_map(n, [=] __device__(int64_t i) {
        p[i] = (invd ? invd[i] * r32[i] : r32[i]) + (float)S->beta * p[i];
      }, hold);
    } else {
      launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
        const float a = (float)S->alpha;
        xi[i] = a * p[i] + xi[i];
        const float ri = r32[i] - a * q[i];
        r32[i] = ri;
        acc[0] += (double)ri * (double)ri;
      }, [=] __device__(const double* t) { rcg_due(S, t[0]); }, hold);
      if (spai) {
        launch_spmv(*M->M, (const float*)r32, EpiRcgPrec{r32, z, S, hold}, s);
      } else {  // a general preconditioner apply is not gated
        M->apply((const float*)r32, z);
        launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
          acc[0] += (double)r32[i] * (double)z[i];
        }, [=] __device__(const double* t) {
          S->beta = t[0] / S->rz;
          S->rz = t[0];
        }, hold);
      }
      launch_map(n, [=] __device__(int64_t i) {
        p[i] = z[i] + (float)S->beta * p[i];
      }, hold);
    }
  };

  restart(true);
  int64_t k = 0;
  if (history) history[0] = rnorm / bnorm;
  // `issued` numbers the queued steps (parity and step stamps); the budget
  // is charged with the steps the device ran (h.iters): steps queued after
  // the device raised `hold` are no-ops and cost nothing
  int64_t ran = 0;
  for (;;) {
    // FP32 steps until an update is due or the budget is spent
    int64_t chunk = 4;
    {
      TraceRange steps_("reliable CG: FP32 steps");
      while (ran < cfg->max_inner) {
        const int64_t e = std::min(cfg->max_inner - ran, chunk);
        for (int64_t j = 0; j < e; ++j, ++issued) step();
        h = st.download();
        ran = h.iters;
        if (h.hold) break;
        chunk = std::min<int64_t>(chunk * 2, 64);
      }
    }
    // reliable update: x64 += xi, FP64 true residual, convergence test
    TraceRange update_("reliable CG: FP64 update");
    launch_map(n, [=] __device__(int64_t i) { x[i] += (double)xi[i]; });
    launch_spmv(*A, (const double*)x, EpiRcgTrue{b, r, S}, s);
    launch_finish<1>(reduce_partials(1), 0, [=] __device__(const double*) {
      S->updates += 1;
      S->rupd = S->rnorm;
      S->converged = S->rnorm / S->bnorm < S->tol ? 1 : 0;
      S->hold = S->converged;
    }, nullptr);
    h = st.download();
    ++k;
    if (history) history[k] = h.rnorm / h.bnorm;
    if (h.converged || k >= cfg->max_outer || h.iters >= cfg->max_inner) break;
    restart(false);
  }
  out->relres = h.rnorm / h.bnorm;
  out->outer_iterations = h.updates;
  out->inner_iterations = h.iters;
  out->converged = h.converged;
  out->status = SPK_OK;
}

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
  if (cfg->inner_method == SPK_IR_CG_RELIABLE) {
    if (h.rnorm / h.bnorm < cfg->tolerance) {
      if (history) history[0] = h.rnorm / h.bnorm;
      out->relres = h.rnorm / h.bnorm;
      out->converged = 1;
      return;
    }
    cg_reliable(A, M, cfg, b, x, r, out, history, h.bnorm, h.rnorm);
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
    TraceRange step_("refinement step");
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
    TraceRange trace_("spk_ir_solve");
    SPK_REQUIRE(A && cfg && d_b && d_x && out, SPK_INVALID_ARGUMENT,
                "null argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "solver requires a square operator");
    SPK_REQUIRE(!M || M->n == A->nrows, SPK_INVALID_ARGUMENT,
                "preconditioner size mismatch");
    SPK_REQUIRE(cfg->max_outer >= 0 && cfg->max_inner >= 0, SPK_INVALID_ARGUMENT,
                "negative iteration budget");
    SPK_REQUIRE(cfg->inner_method >= SPK_IR_GMRES &&
                    cfg->inner_method <= SPK_IR_CG_RELIABLE,
                SPK_INVALID_ARGUMENT, "unknown inner method");
    // an inner mixed-precision solve would reuse this solve's work vectors
    SPK_REQUIRE(!M || M->kind != spk_precond_s::kInnerIr, SPK_INVALID_ARGUMENT,
                "ir_solve: an inner-solve preconditioner cannot be used here");
    ir_solve(A, M, cfg, d_b, d_x, out, history);
    check_preconds(M, nullptr);
  });
}
