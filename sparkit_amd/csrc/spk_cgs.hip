// spk_cgs.hip -- device CGS (cgs.hpp:86-395, Templates Algorithm 7.5) with
// both preconditioner slots composed, computed_iterations counting the two
// matrix-vector products of each step.  Per step:
//   p^ = M p;  q^ = A p^ with <r^, q^> in the SpMV epilogue -> alpha
//   q = u - alpha q^, u^ = u + q                    (one pass)
//   p^ = M u^;  t = A p^ with x += alpha p^, r -= alpha t, |r|^2 in the
//   epilogue -> convergence test
//   rho_new = <r^, r> -> beta;  u = r + beta q, p = u + beta (q + beta p)
// Scalars stay on the device; the host queues steps in chunks.
// Breakdowns the reference leaves undetected stop with SPK_BREAKDOWN:
// <r^, q^> == 0 (a division by zero, :224-226) and rho_new == 0 (the
// reference returns from the step and repeats it with stale directions,
// :206-207).
#include "spk_solver_common.hpp"

namespace spk {
namespace {

struct CgsState {
  double bnorm, tol, resnorm, rho, alpha, beta;
  int64_t iters, steps;
  int32_t done, converged, breakdown, collect;
  double* history;
  double rho_b[2];  // one-kernel hand-off: rho by step parity
};

struct CgsGateNotDone {
  const int* done;
  __device__ bool operator()() const { return !*done; }
};

template <typename T>
struct EpiCgsQParts {  // q^ = A p^, partial <r^, q^> (summed by the next kernel)
  static constexpr bool kReduce = true;
  const T* __restrict__ rh;
  T* __restrict__ qh;
  const int* done;
  __device__ double row(int i, T s) const {
    qh[i] = s;
    return (double)rh[i] * (double)s;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
struct EpiCgsQ {  // q^ = A p^, partial <r^, q^>; finisher: alpha
  static constexpr bool kReduce = true;
  const T* __restrict__ rh;
  T* __restrict__ qh;
  CgsState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    qh[i] = s;
    return (double)rh[i] * (double)s;
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
struct EpiCgsR {  // t = A p^; x += alpha p^; r -= alpha t; partial |r|^2
  static constexpr bool kReduce = true;
  const T* __restrict__ ph;
  T* __restrict__ x;
  T* __restrict__ r;
  CgsState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    const double a = st->alpha;
    x[i] = (T)((double)x[i] + a * (double)ph[i]);
    const T ri = (T)((double)r[i] - a * (double)s);
    r[i] = ri;
    return (double)ri * (double)ri;
  }
  __device__ void finalize(double t) const {
    st->iters += 2;
    st->steps += 1;
    const double rn = sqrt(t);
    st->resnorm = rn;
    if (st->collect) st->history[st->steps - 1] = rn;
    if (rn / st->bnorm < st->tol) {
      st->converged = 1;
      st->done = 1;
    }
  }
};

template <typename T>
void cgs_solve(spk_csr_s* A, spk_precond_s* ML, spk_precond_s* MR,
               const spk_solver_config* cfg, const T* b, T* x, spk_summary* out,
               double* history) {
  const int64_t n = A->nrows;
  const size_t vb = (size_t)(n > 0 ? n : 1) * sizeof(T);
  Workspace& ws = workspace();
  T* r = (T*)ws.get(kWsCgs + 0, vb);
  T* rh = (T*)ws.get(kWsCgs + 1, vb);
  T* p = (T*)ws.get(kWsCgs + 2, vb);
  T* u = (T*)ws.get(kWsCgs + 3, vb);
  T* q = (T*)ws.get(kWsCgs + 4, vb);
  T* qh = (T*)ws.get(kWsCgs + 5, vb);
  T* uh = (T*)ws.get(kWsCgs + 6, vb);
  T* ph = (T*)ws.get(kWsCgs + 7, vb);
  T* tmp = (T*)ws.get(kWsCgs + 8, vb);
  const int64_t max_steps = (cfg->max_iterations + 1) / 2;
  double* dhist = cfg->collect_residuals
                      ? (double*)ws.get(kWsCgs + 9, sizeof(double) *
                                        (size_t)std::max<int64_t>(max_steps, 1))
                      : nullptr;
  DevState<CgsState> st(kWsCgs + 10);
  CgsState h{};
  h.tol = cfg->tolerance;
  h.collect = cfg->collect_residuals ? 1 : 0;
  h.history = dhist;
  st.upload(h);
  CgsState* S = st.d;
  hipStream_t s = stream();
  const int* done = &S->done;

  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)b[i] * (double)b[i];
  }, [=] __device__(const double* tt) { S->bnorm = sqrt(tt[0]); });
  h = st.download();
  if (h.bnorm == 0.0) {  // cgs.hpp:170-174
    launch_map(n, [=] __device__(int64_t i) { x[i] = T(0); });
    *out = spk_summary{0.0, 0, 1, SPK_OK, 0};
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }
  // r = b - A x; r^ = p = u = r; rho = <r^, r> (:176-181, :196-206)
  launch_spmv(*A, (const T*)x, EpiResidual<T>{b, r}, s);
  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    const T ri = r[i];
    rh[i] = ri;
    p[i] = ri;
    u[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }, [=] __device__(const double* tt) {
    S->rho = tt[0];
    S->rho_b[0] = tt[0];
  });

  auto compose = [&](const T* in, T* outv) -> const T* {
    const bool l = !is_identity(ML), rr = !is_identity(MR);
    if (!l && !rr) return in;
    if (l && !rr) { ML->apply(in, outv); return outv; }
    if (!l && rr) { MR->apply(in, outv); return outv; }
    ML->apply(in, tmp);
    MR->apply((const T*)tmp, outv);
    return outv;
  };
  // One-kernel hand-off (the SpMV has few row blocks): no finisher
  // kernels; every reduction's partials are summed by every block of the
  // next kernel.  Per step: q^ = A p^ (partials <r^, q^>); pass 1 (alpha;
  // q, u^); t = A p^' with x, r updates (partials |r|^2); pass 2 (the |r|
  // test; partials rho_new = <r^, r>); pass 3 (beta; u, p).  Block 0 records
  // state no block of the same kernel reads (rho by step parity); a block
  // 0 that ends the solve makes every block of its kernel skip its pass.
  const bool handoff = spmv_partials_count(*A) <= kProMaxPartials &&
                       pro_grid(n) <= kProMaxPartials;
  double* pq[3] = {nullptr, nullptr, nullptr};
  if (handoff) {
    const size_t pb = sizeof(double) * 2 * (size_t)kProMaxPartials;
    for (int i = 0; i < 3; ++i) pq[i] = (double*)ws.get(kWsCgsParts + i, pb);
  }
  // diagonal preconditioner slots (at least one Jacobi): u^' = M u^ and the
  // next p^ = M p are written by the passes that form u^ and p
  // (M_R(M_L(.)) elementwise, the same products as the apply kernels)
  const bool dcomp = is_diagonal(ML) && is_diagonal(MR) &&
                     !(is_identity(ML) && is_identity(MR));
  const T* dml = dcomp ? diag_of<T>(ML) : nullptr;
  const T* dmr = dcomp ? diag_of<T>(MR) : nullptr;
  auto issue_handoff = [&](int64_t k) {
    const int par = (int)(k & 1);
    const T* php = (dcomp && k > 0) ? (const T*)ph : compose(p, ph);
    const int c_q = launch_spmv_partials(*A, php, EpiCgsQParts<T>{rh, qh, done}, pq[0], s);
    // alpha (:224-226); q = u - alpha q^, u^ = u + q (:228-236)
    launch_reduce_pro<1, 0, 0>(
        n, pq[0], c_q, (const double*)nullptr, 0, CgsGateNotDone{done},
        [=] __device__(const double* t, double* sc, bool b0) {
          const double alpha = S->rho_b[par] / t[0];
          sc[0] = alpha;
          sc[1] = t[0] == 0.0 ? 1.0 : 0.0;
          if (b0) {
            if (t[0] == 0.0) {
              S->breakdown = 1;
              S->done = 1;
            } else {
              S->alpha = alpha;
            }
          }
        },
        [=] __device__(int64_t i) { return Pack4<T>{u[i], qh[i], T(0), T(0)}; },
        [=] __device__(int64_t i, const Pack4<T>& e, double*, const double* sc) {
          if (sc[1] != 0.0) return;
          const double qi = (double)e.a - sc[0] * (double)e.b;
          q[i] = (T)qi;
          const T ui = (T)((double)e.a + (double)(T)qi);
          uh[i] = ui;
          if (dcomp) {
            T t = dml ? dml[i] * ui : ui;
            ph[i] = dmr ? dmr[i] * t : t;
          }
        }, (double*)nullptr);
    const T* uhp = dcomp ? (const T*)ph : compose(uh, ph);
    const int c_r = launch_spmv_partials(*A, uhp, EpiCgsR<T>{uhp, x, r, S, done}, pq[1], s);
    // the |r| test (:258-270); rho_new = <r^, r> (:196-206)
    const int c_rho = launch_reduce_pro<1, 0, 1>(
        n, pq[1], c_r, (const double*)nullptr, 0, CgsGateNotDone{done},
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
        [=] __device__(int64_t i) { return Pack4<T>{rh[i], r[i], T(0), T(0)}; },
        [=] __device__(int64_t, const Pack4<T>& e, double* acc, const double* sc) {
          if (sc[0] != 0.0) return;
          acc[0] += (double)e.a * (double)e.b;
        }, pq[2]);
    // beta (:206-209); u = r + beta q, p = u + beta (q + beta p) (:211-222)
    launch_reduce_pro<1, 0, 0>(
        n, pq[2], c_rho, (const double*)nullptr, 0, CgsGateNotDone{done},
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
        [=] __device__(int64_t i) { return Pack4<T>{r[i], q[i], p[i], T(0)}; },
        [=] __device__(int64_t i, const Pack4<T>& e, double*, const double* sc) {
          if (sc[1] != 0.0) return;
          const double be = sc[0];
          const double qi = (double)e.b;
          const T ui = (T)((double)e.a + be * qi);
          u[i] = ui;
          const T pi = (T)((double)ui + be * (qi + be * (double)e.c));
          p[i] = pi;
          if (dcomp) {
            T t = dml ? dml[i] * pi : pi;
            ph[i] = dmr ? dmr[i] * t : t;
          }
        }, (double*)nullptr);
  };
  auto issue = [&](int64_t k) {
    if (handoff) {
      issue_handoff(k);
      return;
    }
    const T* php = compose(p, ph);
    launch_spmv(*A, php, EpiCgsQ<T>{rh, qh, S, done}, s);
    launch_map(n, [=] __device__(int64_t i) {
      const double qi = (double)u[i] - S->alpha * (double)qh[i];
      q[i] = (T)qi;
      uh[i] = (T)((double)u[i] + (double)(T)qi);
    }, done);
    const T* uhp = compose(uh, ph);
    launch_spmv(*A, uhp, EpiCgsR<T>{uhp, x, r, S, done}, s);
    launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
      acc[0] += (double)rh[i] * (double)r[i];
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
      const double qi = (double)q[i];
      const T ui = (T)((double)r[i] + be * qi);
      u[i] = ui;
      p[i] = (T)((double)ui + be * (qi + be * (double)p[i]));
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

extern "C" spk_status spk_cgs_solve(spk_csr A, spk_precond M_left,
                                    spk_precond M_right,
                                    const spk_solver_config* cfg, const void* d_b,
                                    void* d_x, spk_summary* summary,
                                    double* history) {
  return guarded([&] {
    TraceRange trace_("spk_cgs_solve");
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(summary, SPK_INVALID_ARGUMENT, "null summary");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      cgs_solve<float>(A, M_left, M_right, cfg, (const float*)d_b, (float*)d_x,
                       summary, history);
    else
      cgs_solve<double>(A, M_left, M_right, cfg, (const double*)d_b, (double*)d_x,
                        summary, history);
    check_preconds(M_left, M_right);
  });
}
