// spk_bicgstab.hip -- device BiCGSTAB (bicgstab.hpp:86-361): van der Vorst's
// recurrence with both preconditioner slots composed on p and s, the
// half-step convergence test on |s|, and computed_iterations counting
// matrix-vector products.  Scalars stay on the device; the host queues
// steps in chunks and polls the done flag.
//
// Breakdowns the reference leaves undetected stop the solve with status
// SPK_BREAKDOWN: rho_new == 0 (the reference returns without consuming
// budget and loops forever, bicgstab.hpp:203-204) and omega == 0 (it would
// divide by omega in the next beta, :206, :222-223).  On omega == 0 the
// step's valid half (x += alpha p^, r = s) is kept before stopping, so the
// returned x satisfies the reported residual norm |s|.
#include "spk_solver_common.hpp"

namespace spk {
namespace {

struct BiState {
  double bnorm, tol, resnorm;
  double rho, rho_new, alpha, omega, beta;
  int64_t iters;
  int64_t half_step;  // step index whose half-step test converged, else -1
  int32_t done, converged, breakdown, collect;
  double* history;
  int64_t k2_step;  // one-kernel hand-off: the step whose s pass ran
};

struct BiGateNotDone {
  const int* done;
  __device__ bool operator()() const { return !*done; }
};
struct BiGateStep {
  const BiState* st;
  int64_t step;
  __device__ bool operator()() const { return st->k2_step == step; }
};

template <typename T>
struct EpiBiVParts {  // v = A p^, partial <r^, v> (summed by the next kernel)
  static constexpr bool kReduce = true;
  const T* __restrict__ rh;
  T* __restrict__ v;
  const int* done;
  __device__ double row(int i, T s) const {
    v[i] = s;
    return (double)rh[i] * (double)s;
  }
  __device__ void finalize(double) const {}
};

// the |r| test that closes a step (bicgstab.hpp:310-324), thread 0 of block 0
__device__ inline void bi_r_test(BiState* S, double rr) {
  const double rn = sqrt(rr);
  S->resnorm = rn;
  if (S->collect) S->history[S->iters - 1] = rn;
  if (rn / S->bnorm < S->tol) {
    S->converged = 1;
    S->done = 1;
  }
}

template <typename T>
struct EpiBiV {  // v = A p^, partial <r^, v>; last block: alpha
  static constexpr bool kReduce = true;
  const T* __restrict__ rh;
  T* __restrict__ v;
  BiState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    v[i] = s;
    return (double)rh[i] * (double)s;
  }
  __device__ void finalize(double t) const {
    st->iters += 1;
    st->alpha = st->rho_new / t;
  }
};

template <typename T>
void bicgstab_solve(spk_csr_s* A, spk_precond_s* ML, spk_precond_s* MR,
                    const spk_solver_config* cfg, const T* b, T* x,
                    spk_summary* out, double* history) {
  const int64_t n = A->nrows;
  const size_t vb = (size_t)(n > 0 ? n : 1) * sizeof(T);
  Workspace& ws = workspace();
  T* r = (T*)ws.get(kWsBicgstab + 0, vb);
  T* rh = (T*)ws.get(kWsBicgstab + 1, vb);
  T* p = (T*)ws.get(kWsBicgstab + 2, vb);
  T* v = (T*)ws.get(kWsBicgstab + 3, vb);
  T* sv = (T*)ws.get(kWsBicgstab + 4, vb);
  T* t = (T*)ws.get(kWsBicgstab + 5, vb);
  T* ph = (T*)ws.get(kWsBicgstab + 6, vb);
  T* sh = (T*)ws.get(kWsBicgstab + 7, vb);
  T* tmp = (T*)ws.get(kWsBicgstab + 8, vb);
  double* dhist = cfg->collect_residuals
                      ? (double*)ws.get(kWsBicgstab + 9, sizeof(double) *
                                        (size_t)(std::max<int64_t>(cfg->max_iterations, 1) + 1))
                      : nullptr;
  DevState<BiState> st(kWsBicgstab + 10);
  BiState h{};
  h.tol = cfg->tolerance;
  h.collect = cfg->collect_residuals ? 1 : 0;
  h.history = dhist;
  h.half_step = -1;
  h.rho = h.alpha = h.omega = 1.0;
  st.upload(h);
  BiState* S = st.d;
  hipStream_t s = stream();
  const int* done = &S->done;

  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)b[i] * (double)b[i];
  }, [=] __device__(const double* tt) { S->bnorm = sqrt(tt[0]); });
  h = st.download();
  if (h.bnorm == 0.0) {  // bicgstab.hpp:160-165
    launch_map(n, [=] __device__(int64_t i) { x[i] = T(0); });
    *out = spk_summary{0.0, 0, 1, SPK_OK, 0};
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }
  // r = b - A x; r^ = r; v = p = 0 (:167-174)
  launch_spmv(*A, (const T*)x, EpiResidual<T>{b, r}, s);
  launch_map(n, [=] __device__(int64_t i) {
    rh[i] = r[i];
    v[i] = T(0);
    p[i] = T(0);
  });

  // M_R^-1 (M_L^-1 v) into out (the reference's tmp + second slot); with
  // both slots the identity the input itself is used
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
  // kernels.  Each reduction's partials are summed by every block of the
  // next kernel, which derives the scalar it needs; block 0 records state
  // that no block of the same kernel reads.  Per step:
  //   K1  prologue: the previous step's |r| test, rho_new = <r^, r>, beta;
  //       p = r + beta (p - omega v)
  //   S1  v = A p^, partials <r^, v>
  //   K2  prologue: alpha; s = r - alpha v, partials |s|^2
  //   S2  t = A s^
  //   K4  prologue: the half-step |s| test -- on exit x += alpha p^ --;
  //       otherwise partials <t,t>, <t,s>  (gated on K2's step stamp, as
  //       block 0 may set done)
  //   K5  prologue: omega; x += alpha p^ + omega s^, r = s - omega t,
  //       partials |r|^2, <r^, r> for the next K1
  // A step's |r| test runs in the next step's K1, or in a tail kernel
  // after the last step.
  const bool handoff = spmv_partials_count(*A) <= kProMaxPartials &&
                       pro_grid(n) <= kProMaxPartials;
  double* pq[4] = {nullptr, nullptr, nullptr, nullptr};
  int c_v = 0, c_s = 0, c_t = 0, c_r = 0;
  if (handoff) {
    const size_t pb = sizeof(double) * 2 * (size_t)kProMaxPartials;
    for (int i = 0; i < 4; ++i) pq[i] = (double*)ws.get(kWsBicgstab + 11 + i, pb);
    // <r^, r> and |r|^2 of r0 for the first K1 (r^ = r)
    c_r = launch_reduce_pro<1, 0, 2>(
        n, pq[3], 0, (const double*)nullptr, 0, BiGateNotDone{done},
        [=] __device__(const double*, double*, bool) {},
        [=] __device__(int64_t i) { return Pack4<T>{r[i], T(0), T(0), T(0)}; },
        [=] __device__(int64_t, const Pack4<T>& e, double* acc, const double*) {
          const double ri = (double)e.a;
          acc[0] += ri * ri;
          acc[1] += ri * ri;
        }, pq[3]);
  }
  // diagonal preconditioner slots (at least one Jacobi): p^ and s^ are
  // written by the passes that form p and s (M_R(M_L(.)) elementwise, the
  // same products as the separate apply kernels)
  const bool dcomp = is_diagonal(ML) && is_diagonal(MR) &&
                     !(is_identity(ML) && is_identity(MR));
  const T* dml = dcomp ? diag_of<T>(ML) : nullptr;
  const T* dmr = dcomp ? diag_of<T>(MR) : nullptr;
  auto issue_handoff = [&](int64_t k) {
    const int64_t sid = k + 1;
    // K1 (:203-206, :232-238; the previous step's :310-324)
    const bool test_r = k > 0;
    launch_reduce_pro<2, 0, 0>(
        n, pq[3], c_r, (const double*)nullptr, 0, BiGateNotDone{done},
        [=] __device__(const double* t, double* sc, bool b0) {
          const double rn2 = t[0], rho_new = t[1];
          bool stop = S->breakdown != 0;  // omega == 0 in the last step
          if (test_r && sqrt(rn2) / S->bnorm < S->tol) stop = true;
          const double beta = (rho_new / S->rho) * (S->alpha / S->omega);
          sc[0] = beta;
          sc[1] = S->omega;
          sc[2] = (stop || rho_new == 0.0) ? 1.0 : 0.0;
          if (b0) {
            if (test_r) bi_r_test(S, rn2);
            if (S->breakdown) S->done = 1;
            if (!stop) {
              S->rho_new = rho_new;
              if (rho_new == 0.0) {
                S->breakdown = 1;
                S->done = 1;
              } else {
                S->beta = beta;
              }
            }
          }
        },
        [=] __device__(int64_t i) { return Pack4<T>{r[i], p[i], v[i], T(0)}; },
        [=] __device__(int64_t i, const Pack4<T>& e, double*, const double* sc) {
          if (sc[2] != 0.0) return;
          const T pi = (T)((double)e.a + sc[0] * ((double)e.b - sc[1] * (double)e.c));
          p[i] = pi;
          if (dcomp) {
            T t = dml ? dml[i] * pi : pi;
            ph[i] = dmr ? dmr[i] * t : t;
          }
        }, (double*)nullptr);
    // S1 (:240-250)
    const T* php = dcomp ? (const T*)ph : compose(p, ph);
    c_v = launch_spmv_partials(*A, php, EpiBiVParts<T>{rh, v, done}, pq[0], s);
    // K2 (:211, :252-258)
    c_s = launch_reduce_pro<1, 0, 1>(
        n, pq[0], c_v, (const double*)nullptr, 0, BiGateNotDone{done},
        [=] __device__(const double* t, double* sc, bool b0) {
          const double alpha = S->rho_new / t[0];
          sc[0] = alpha;
          if (b0) {
            S->iters += 1;
            S->alpha = alpha;
            S->k2_step = sid;
          }
        },
        [=] __device__(int64_t i) { return Pack4<T>{r[i], v[i], T(0), T(0)}; },
        [=] __device__(int64_t i, const Pack4<T>& e, double* acc, const double* sc) {
          const T si = (T)((double)e.a - sc[0] * (double)e.b);
          sv[i] = si;
          if (dcomp) {
            T t = dml ? dml[i] * si : si;
            sh[i] = dmr ? dmr[i] * t : t;
          }
          acc[0] += (double)si * (double)si;
        }, pq[1]);
    // S2 (:280-290)
    const T* shp = dcomp ? (const T*)sh : compose(sv, sh);
    launch_spmv(*A, shp, EpiStore<T>{t, done}, s);
    // K4 (:259-278, :292-297)
    c_t = launch_reduce_pro<1, 0, 2>(
        n, pq[1], c_s, (const double*)nullptr, 0, BiGateStep{S, sid},
        [=] __device__(const double* tt, double* sc, bool b0) {
          const double sn = sqrt(tt[0]);
          const bool conv = sn / S->bnorm < S->tol;
          sc[0] = S->alpha;
          sc[1] = conv ? 1.0 : 0.0;
          if (b0) {
            S->resnorm = sn;
            if (S->collect) S->history[S->iters - 1] = sn;
            if (conv) {
              S->half_step = sid - 1;
              S->converged = 1;
              S->done = 1;
            }
          }
        },
        [=] __device__(int64_t i) { return Pack4<T>{t[i], sv[i], x[i], php[i]}; },
        [=] __device__(int64_t i, const Pack4<T>& e, double* acc, const double* sc) {
          if (sc[1] != 0.0) {  // half-step exit (:271-278)
            x[i] = (T)((double)e.c + sc[0] * (double)e.d);
            return;
          }
          const double ti = (double)e.a;
          acc[0] += ti * ti;
          acc[1] += ti * (double)e.b;
        }, pq[2]);
    // K5 (:292-324)
    c_r = launch_reduce_pro<2, 0, 2>(
        n, pq[2], c_t, (const double*)nullptr, 0, BiGateNotDone{done},
        [=] __device__(const double* tt, double* sc, bool b0) {
          const double om = tt[0] == 0.0 ? 0.0 : tt[1] / tt[0];
          sc[0] = S->alpha;
          sc[1] = om;
          if (b0) {
            S->iters += 1;
            S->omega = om;
            S->rho = S->rho_new;
            if (om == 0.0) S->breakdown = 1;
          }
        },
        [=] __device__(int64_t i) { return Pack8<T>{x[i], php[i], shp[i], sv[i], t[i],
                                                    rh[i], T(0), T(0)}; },
        [=] __device__(int64_t i, const Pack8<T>& e, double* acc, const double* sc) {
          // omega == 0 (breakdown) makes this the half step: x += alpha p^,
          // r = s; the next K1 (or the tail) tests |r| and ends the solve
          x[i] = (T)((double)e.a + (sc[0] * (double)e.b + sc[1] * (double)e.c));
          const T ri = (T)((double)e.d - sc[1] * (double)e.e);
          r[i] = ri;
          acc[0] += (double)ri * (double)ri;
          acc[1] += (double)e.f * (double)ri;
        }, pq[3]);
  };

  auto issue = [&](int64_t k) {
    if (handoff) {
      issue_handoff(k);
      return;
    }
    // rho_new = <r^, r>; beta (:203-206)
    launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
      acc[0] += (double)rh[i] * (double)r[i];
    }, [=] __device__(const double* tt) {
      double rn = tt[0];
      S->rho_new = rn;
      if (rn == 0.0) {
        S->breakdown = 1;
        S->done = 1;
        return;
      }
      S->beta = (rn / S->rho) * (S->alpha / S->omega);
    }, done);
    // p = r + beta (p - omega v) (:232-238)
    launch_map(n, [=] __device__(int64_t i) {
      p[i] = (T)((double)r[i] + S->beta * ((double)p[i] - S->omega * (double)v[i]));
    }, done);
    // p^ = M_R M_L p (:240-244); v = A p^ (:246-250); alpha (:211)
    const T* php = compose(p, ph);
    launch_spmv(*A, php, EpiBiV<T>{rh, v, S, done}, s);
    // s = r - alpha v; half-step test on |s| (:252-278)
    launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
      T si = (T)((double)r[i] - S->alpha * (double)v[i]);
      sv[i] = si;
      acc[0] += (double)si * (double)si;
    }, [=] __device__(const double* tt) {
      double sn = sqrt(tt[0]);
      S->resnorm = sn;
      if (S->collect) S->history[S->iters - 1] = sn;
      if (sn / S->bnorm < S->tol) {
        S->half_step = k;
        S->converged = 1;
        S->done = 1;
      }
    }, done);
    // half-step exit: x += alpha p^ (:271-278); runs only for this step
    launch_map(n, [=] __device__(int64_t i) {
      if (S->half_step == k) x[i] = (T)((double)x[i] + S->alpha * (double)php[i]);
    });
    // s^ = M_R M_L s; t = A s^ (:280-290)
    const T* shp = compose(sv, sh);
    launch_spmv(*A, shp, EpiStore<T>{t, done}, s);
    // omega = <t,s>/<t,t> (:292-297)
    launch_reduce<2>(n, [=] __device__(int64_t i, double* acc) {
      double ti = (double)t[i];
      acc[0] += ti * ti;
      acc[1] += ti * (double)sv[i];
    }, [=] __device__(const double* tt) {
      S->iters += 1;
      double om = tt[0] == 0.0 ? 0.0 : tt[1] / tt[0];
      S->omega = om;
      // omega == 0: the update below still runs (with omega = 0 it is the
      // half step x += alpha p^, r = s) and its finisher ends the solve
      if (om == 0.0) S->breakdown = 1;
    }, done);
    // x += alpha p^ + omega s^; r = s - omega t; |r| (:299-324)
    launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
      x[i] = (T)((double)x[i] + (S->alpha * (double)php[i] + S->omega * (double)shp[i]));
      T ri = (T)((double)sv[i] - S->omega * (double)t[i]);
      r[i] = ri;
      acc[0] += (double)ri * (double)ri;
    }, [=] __device__(const double* tt) {
      S->rho = S->rho_new;
      double rn = sqrt(tt[0]);
      S->resnorm = rn;
      if (S->collect) S->history[S->iters - 1] = rn;
      if (rn / S->bnorm < S->tol) {
        S->converged = 1;
        S->done = 1;
      }
      if (S->breakdown) S->done = 1;
    }, done);
  };
  // each step consumes two matvecs; steps run while budget remains (:176)
  const int64_t steps = (cfg->max_iterations + 1) / 2;
  auto is_done = [&]() { return st.download().done != 0; };
  chunked_loop(steps, issue, is_done, 2, 32);
  if (handoff && steps > 0) {
    // the last step's |r| test (:310-324)
    launch_reduce_pro<2, 0, 0>(
        1, pq[3], c_r, (const double*)nullptr, 0, BiGateNotDone{done},
        [=] __device__(const double* t, double*, bool b0) {
          if (b0) {
            bi_r_test(S, t[0]);
            if (S->breakdown) S->done = 1;
          }
        },
        [=] __device__(int64_t) { return Pack4<T>{T(0), T(0), T(0), T(0)}; },
        [=] __device__(int64_t, const Pack4<T>&, double*, const double*) {},
        (double*)nullptr);
  }

  h = st.download();
  out->residual_norm = h.resnorm;
  out->computed_iterations = h.iters;
  out->converged = h.converged;
  out->status = h.breakdown ? SPK_BREAKDOWN : SPK_OK;
  // one history entry per residual test: per matvec pair, the |s| test and
  // (unless it converged) the |r| test
  out->history_len = 0;
  if (h.collect) out->history_len = h.iters;
  if (history && h.collect && h.iters > 0) {
    SPK_HIP_TRY(hipMemcpyAsync(history, dhist, h.iters * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
  }
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_bicgstab_solve(spk_csr A, spk_precond M_left,
                                         spk_precond M_right,
                                         const spk_solver_config* cfg,
                                         const void* d_b, void* d_x,
                                         spk_summary* summary,
                                         double* history) {
  return guarded([&] {
    TraceRange trace_("spk_bicgstab_solve");
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(summary, SPK_INVALID_ARGUMENT, "null summary");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      bicgstab_solve<float>(A, M_left, M_right, cfg, (const float*)d_b,
                            (float*)d_x, summary, history);
    else
      bicgstab_solve<double>(A, M_left, M_right, cfg, (const double*)d_b,
                             (double*)d_x, summary, history);
    check_preconds(M_left, M_right);
  });
}
