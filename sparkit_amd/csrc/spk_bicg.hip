// spk_bicg.hip -- device BiCG (bicg.hpp:86-397) with a transpose operator
// given as the store of A^T (a store's SpMV adds each row in storage order,
// which for A^T is multiply_transpose's order, sparse_blas.hpp:36-52).
// Primal preconditioning applies M_L then M_R, the adjoint M_R then M_L
// (bicg.hpp:215-222; both assumed symmetric).  Per step:
//   q = A p with <p~, q> in the epilogue -> alpha
//   x += alpha p, r -= alpha q, |r|^2 -> convergence test (two matvecs
//   counted, :242)
//   q~ = A^T p~ with r~ -= alpha q~ in the epilogue
//   z = M r, z~ = M^T r~;  rho_new = <r~, z> -> beta
//   p = z + beta p, p~ = z~ + beta p~
// rho_new == 0 (the reference returns and repeats the step, :246-247) and
// <p~, q> == 0 stop with SPK_BREAKDOWN.
#include "spk_solver_common.hpp"

namespace spk {
namespace {

struct BicgState {
  double bnorm, tol, resnorm, rho, alpha, beta;
  int64_t iters, steps;
  int32_t done, converged, breakdown, collect;
  double* history;
  double rho_b[2];  // one-kernel hand-off: rho by step parity
};

struct BicgGateNotDone {
  const int* done;
  __device__ bool operator()() const { return !*done; }
};

template <typename T>
struct EpiBicgShadowZ {  // q~ = A^T p~; r~ -= alpha q~; z~ = M_L(M_R(r~)) diagonal
  static constexpr bool kReduce = false;
  T* __restrict__ rt;
  T* __restrict__ zt;
  const T* __restrict__ dl;
  const T* __restrict__ dr;
  const BicgState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    const T ri = (T)((double)rt[i] - st->alpha * (double)s);
    rt[i] = ri;
    T t = dr ? dr[i] * ri : ri;
    zt[i] = dl ? dl[i] * t : t;
    return 0.0;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
struct EpiBicgQParts {  // q = A p, partial <p~, q> (summed by the next kernel)
  static constexpr bool kReduce = true;
  const T* __restrict__ pt;
  T* __restrict__ q;
  const int* done;
  __device__ double row(int i, T s) const {
    q[i] = s;
    return (double)pt[i] * (double)s;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
struct EpiBicgQ {  // q = A p, partial <p~, q>; finisher: alpha
  static constexpr bool kReduce = true;
  const T* __restrict__ pt;
  T* __restrict__ q;
  BicgState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    q[i] = s;
    return (double)pt[i] * (double)s;
  }
  __device__ void finalize(double t) const {
    if (t == 0.0) {
      st->breakdown = 1;
      st->done = 1;
      return;
    }
    st->alpha = st->rho / t;
  }
};

template <typename T>
struct EpiBicgShadow {  // q~ = A^T p~; r~ -= alpha q~
  static constexpr bool kReduce = false;
  T* __restrict__ rt;
  const BicgState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    rt[i] = (T)((double)rt[i] - st->alpha * (double)s);
    return 0.0;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
void bicg_solve(spk_csr_s* A, spk_csr_s* At, spk_precond_s* ML, spk_precond_s* MR,
                const spk_solver_config* cfg, const T* b, T* x, spk_summary* out,
                double* history) {
  const int64_t n = A->nrows;
  const size_t vb = (size_t)(n > 0 ? n : 1) * sizeof(T);
  Workspace& ws = workspace();
  T* r = (T*)ws.get(kWsBicg + 0, vb);
  T* rt = (T*)ws.get(kWsBicg + 1, vb);
  T* z = (T*)ws.get(kWsBicg + 2, vb);
  T* zt = (T*)ws.get(kWsBicg + 3, vb);
  T* p = (T*)ws.get(kWsBicg + 4, vb);
  T* pt = (T*)ws.get(kWsBicg + 5, vb);
  T* q = (T*)ws.get(kWsBicg + 6, vb);
  T* tmp = (T*)ws.get(kWsBicg + 7, vb);
  const int64_t max_steps = (cfg->max_iterations + 1) / 2;
  double* dhist = cfg->collect_residuals
                      ? (double*)ws.get(kWsBicg + 8, sizeof(double) *
                                        (size_t)std::max<int64_t>(max_steps, 1))
                      : nullptr;
  DevState<BicgState> st(kWsBicg + 9);
  BicgState h{};
  h.tol = cfg->tolerance;
  h.collect = cfg->collect_residuals ? 1 : 0;
  h.history = dhist;
  st.upload(h);
  BicgState* S = st.d;
  hipStream_t s = stream();
  const int* done = &S->done;

  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)b[i] * (double)b[i];
  }, [=] __device__(const double* tt) { S->bnorm = sqrt(tt[0]); });
  h = st.download();
  if (h.bnorm == 0.0) {  // bicg.hpp:186-190
    launch_map(n, [=] __device__(int64_t i) { x[i] = T(0); });
    *out = spk_summary{0.0, 0, 1, SPK_OK, 0};
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }
  // z = M_R M_L v (primal), M_L M_R v (adjoint); identity slots pass through
  auto precond = [&](const T* in, T* outv, bool adjoint) {
    spk_precond_s* first = adjoint ? MR : ML;
    spk_precond_s* second = adjoint ? ML : MR;
    const bool f = !is_identity(first), g = !is_identity(second);
    if (!f && !g) {
      launch_map(n, [=] __device__(int64_t i) { outv[i] = in[i]; });
    } else if (f && !g) {
      first->apply(in, outv);
    } else if (!f && g) {
      second->apply(in, outv);
    } else {
      first->apply(in, tmp);
      second->apply((const T*)tmp, outv);
    }
  };
  // r = b - A x; r~ = r; z, z~; p = z, p~ = z~; rho = <r~, z> (:192-200)
  launch_spmv(*A, (const T*)x, EpiResidual<T>{b, r}, s);
  launch_map(n, [=] __device__(int64_t i) { rt[i] = r[i]; });
  precond(r, z, false);
  precond(rt, zt, true);
  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    p[i] = z[i];
    pt[i] = zt[i];
    acc[0] += (double)rt[i] * (double)z[i];
  }, [=] __device__(const double* tt) {
    S->rho = tt[0];
    S->rho_b[0] = tt[0];
  });

  // One-kernel hand-off (the SpMV has few row blocks): no finisher
  // kernels; every reduction's partials are summed by every block of the
  // next kernel.  Per step: q = A p (partials <p~, q>); pass 1 (alpha;
  // x, r; partials |r|^2); r~ -= alpha A^T p~; z, z~; pass 2 (the |r| test;
  // partials rho_new = <r~, z>); pass 3 (beta; p, p~).  Block 0 records
  // state no block of the same kernel reads (rho by step parity); a block
  // 0 that ends the solve makes every block of its kernel skip its pass.
  const bool handoff = spmv_partials_count(*A) <= kProMaxPartials &&
                       pro_grid(n) <= kProMaxPartials;
  double* pq[3] = {nullptr, nullptr, nullptr};
  if (handoff) {
    const size_t pb = sizeof(double) * 2 * (size_t)kProMaxPartials;
    for (int i = 0; i < 3; ++i) pq[i] = (double*)ws.get(kWsBicgParts + i, pb);
  }
  // diagonal preconditioner slots: z = M_R(M_L(r)) and z~ = M_L(M_R(r~))
  // are written by the pass that forms r and by the shadow SpMV's epilogue
  // (the same products as the apply kernels; a copy for identity slots)
  const bool dcomp = is_diagonal(ML) && is_diagonal(MR);
  const T* dml = dcomp ? diag_of<T>(ML) : nullptr;
  const T* dmr = dcomp ? diag_of<T>(MR) : nullptr;
  auto issue_handoff = [&](int64_t k) {
    const int par = (int)(k & 1);
    const int c_q = launch_spmv_partials(*A, (const T*)p, EpiBicgQParts<T>{pt, q, done},
                                         pq[0], s);
    // alpha (:214-217); x += alpha p, r -= alpha q (:219-224)
    const int c_r = launch_reduce_pro<1, 0, 1>(
        n, pq[0], c_q, (const double*)nullptr, 0, BicgGateNotDone{done},
        [=] __device__(const double* t, double* sc, bool b0) {
          sc[0] = S->rho_b[par] / t[0];
          sc[1] = t[0] == 0.0 ? 1.0 : 0.0;
          if (b0) {
            if (t[0] == 0.0) {
              S->breakdown = 1;
              S->done = 1;
            } else {
              S->alpha = sc[0];
            }
          }
        },
        [=] __device__(int64_t i) { return Pack4<T>{x[i], p[i], r[i], q[i]}; },
        [=] __device__(int64_t i, const Pack4<T>& e, double* acc, const double* sc) {
          if (sc[1] != 0.0) return;
          const double a = sc[0];
          x[i] = (T)((double)e.a + a * (double)e.b);
          const T ri = (T)((double)e.c - a * (double)e.d);
          r[i] = ri;
          if (dcomp) {
            T t = dml ? dml[i] * ri : ri;
            z[i] = dmr ? dmr[i] * t : t;
          }
          acc[0] += (double)ri * (double)ri;
        }, pq[1]);
    if (dcomp) {
      launch_spmv(*At, (const T*)pt, EpiBicgShadowZ<T>{rt, zt, dml, dmr, S, done}, s);
    } else {
      launch_spmv(*At, (const T*)pt, EpiBicgShadow<T>{rt, S, done}, s);
      precond(r, z, false);
      precond(rt, zt, true);
    }
    // the |r| test (:226-238); rho_new = <r~, z> (:240-246)
    const int c_rho = launch_reduce_pro<1, 0, 1>(
        n, pq[1], c_r, (const double*)nullptr, 0, BicgGateNotDone{done},
        [=] __device__(const double* t, double* sc, bool b0) {
          const double rn = sqrt(t[0]);
          const bool conv = rn / S->bnorm < S->tol;
          sc[0] = conv ? 1.0 : 0.0;
          if (b0) {
            S->iters += 2;
            S->steps += 1;
            S->resnorm = rn;
            if (S->collect) S->history[S->steps - 1] = rn;
            if (conv) {
              S->converged = 1;
              S->done = 1;
            }
          }
        },
        [=] __device__(int64_t i) { return Pack4<T>{rt[i], z[i], T(0), T(0)}; },
        [=] __device__(int64_t, const Pack4<T>& e, double* acc, const double* sc) {
          if (sc[0] != 0.0) return;
          acc[0] += (double)e.a * (double)e.b;
        }, pq[2]);
    // beta (:246-250); p = z + beta p, p~ = z~ + beta p~ (:252-258)
    launch_reduce_pro<1, 0, 0>(
        n, pq[2], c_rho, (const double*)nullptr, 0, BicgGateNotDone{done},
        [=] __device__(const double* t, double* sc, bool b0) {
          const double rn = t[0];
          sc[0] = rn / S->rho_b[par];
          sc[1] = rn == 0.0 ? 1.0 : 0.0;
          if (b0) {
            if (rn == 0.0) {
              S->breakdown = 1;
              S->done = 1;
            } else {
              S->beta = sc[0];
              S->rho = rn;
              S->rho_b[par ^ 1] = rn;
            }
          }
        },
        [=] __device__(int64_t i) { return Pack4<T>{z[i], p[i], zt[i], pt[i]}; },
        [=] __device__(int64_t i, const Pack4<T>& e, double*, const double* sc) {
          if (sc[1] != 0.0) return;
          const double be = sc[0];
          p[i] = (T)((double)e.a + be * (double)e.b);
          pt[i] = (T)((double)e.c + be * (double)e.d);
        }, (double*)nullptr);
  };

  auto issue = [&](int64_t k) {
    if (handoff) {
      issue_handoff(k);
      return;
    }
    launch_spmv(*A, (const T*)p, EpiBicgQ<T>{pt, q, S, done}, s);
    launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
      const double a = S->alpha;
      x[i] = (T)((double)x[i] + a * (double)p[i]);
      const T ri = (T)((double)r[i] - a * (double)q[i]);
      r[i] = ri;
      acc[0] += (double)ri * (double)ri;
    }, [=] __device__(const double* tt) {
      S->iters += 2;
      S->steps += 1;
      const double rn = sqrt(tt[0]);
      S->resnorm = rn;
      if (S->collect) S->history[S->steps - 1] = rn;
      if (rn / S->bnorm < S->tol) {
        S->converged = 1;
        S->done = 1;
      }
    }, done);
    launch_spmv(*At, (const T*)pt, EpiBicgShadow<T>{rt, S, done}, s);
    precond(r, z, false);
    precond(rt, zt, true);
    launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
      acc[0] += (double)rt[i] * (double)z[i];
    }, [=] __device__(const double* tt) {
      const double rn = tt[0];
      if (rn == 0.0) {
        S->breakdown = 1;
        S->done = 1;
        return;
      }
      S->beta = rn / S->rho;
      S->rho = rn;
    }, done);
    launch_map(n, [=] __device__(int64_t i) {
      const double be = S->beta;
      p[i] = (T)((double)z[i] + be * (double)p[i]);
      pt[i] = (T)((double)zt[i] + be * (double)pt[i]);
    }, done);
  };
  auto is_done = [&]() { return st.download().done != 0; };
  chunked_loop(max_steps, issue, is_done, 2, 32);

  h = st.download();
  out->residual_norm = h.resnorm;
  out->computed_iterations = h.iters;
  out->converged = h.converged;
  out->status = h.breakdown ? SPK_BREAKDOWN : SPK_OK;
  out->history_len = h.collect ? h.steps : 0;
  if (history && h.collect && h.steps > 0) {
    SPK_HIP_TRY(hipMemcpyAsync(history, dhist, h.steps * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
  }
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_bicg_solve(spk_csr A, spk_csr At, spk_precond M_left,
                                     spk_precond M_right,
                                     const spk_solver_config* cfg, const void* d_b,
                                     void* d_x, spk_summary* summary,
                                     double* history) {
  return guarded([&] {
    TraceRange trace_("spk_bicg_solve");
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(At && summary, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(At->nrows == A->ncols && At->ncols == A->nrows && At->nnz == A->nnz,
                SPK_INVALID_ARGUMENT, "At must be the transpose of A");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      bicg_solve<float>(A, At, M_left, M_right, cfg, (const float*)d_b, (float*)d_x,
                        summary, history);
    else
      bicg_solve<double>(A, At, M_left, M_right, cfg, (const double*)d_b,
                         (double*)d_x, summary, history);
    check_preconds(M_left, M_right);
  });
}
