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
  }, [=] __device__(const double* tt) { S->rho = tt[0]; });

  auto compose = [&](const T* in, T* outv) -> const T* {
    const bool l = !is_identity(ML), rr = !is_identity(MR);
    if (!l && !rr) return in;
    if (l && !rr) { ML->apply(in, outv); return outv; }
    if (!l && rr) { MR->apply(in, outv); return outv; }
    ML->apply(in, tmp);
    MR->apply((const T*)tmp, outv);
    return outv;
  };
  auto issue = [&](int64_t) {
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
