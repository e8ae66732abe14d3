// spk_qmr.hip -- device QMR without look-ahead (qmr.hpp:86-440): two-sided
// Lanczos with the transpose operator given as the store of A^T, the
// quasi-minimal residual smoothing kept in device scalars.  Primal chain
// z = M_R(M_L(A p)), adjoint chain w~ = A^T M_L(M_R(q)).  Per step:
//   K1  delta = <w, v>;  finisher: breakdown test, the direction factors
//   K2  p = v - f1 p, q = w - f2 q (copies on the first step); for
//       diagonal preconditioners also u = M_L(M_R(q))
//   K3  z = M_R(M_L(A p)) in the SpMV epilogue, <q, z>;  finisher: epsilon,
//       beta = epsilon / delta
//   K4  v~ = z - beta v, |v~|^2;  finisher: rho_new
//   K5  w~ = A^T u - beta w in the epilogue, |w~|^2;  finisher: xi_new,
//       theta, gamma, tau, eta, the residual estimate |tau| and the
//       convergence test (two matvecs counted, :251)
//   K6  d = eta p + (theta_old gamma)^2 d; x += d; v = v~ / rho_new,
//       w = w~ / xi_new (gated on the step count)
// delta == 0 (the reference returns without progress and repeats the step
// forever, :242) and rho_new or xi_new == 0 without convergence (:258)
// stop with SPK_BREAKDOWN.
#include "spk_solver_common.hpp"

namespace spk {
namespace {

// the smoothing state one step hands to the next
struct QmrSmooth {
  double rho, xi, gamma, eta, theta, tau;
};

struct QmrState {
  double bnorm, tol, resnorm, rho, xi, gamma, eta, theta, eps, beta, tau;
  double delta, f1, f2, coeff, rho_new;
  int64_t iters, steps;
  int32_t done, converged, breakdown, collect;
  double* history;
  // one-kernel hand-off: the smoothing state by step parity, and the step
  // whose v~ pass ran (gates the update pass)
  QmrSmooth q[2];
  int64_t k4_step;
};

// qmr_smooth's arithmetic as a pure function of the previous step's state
// (qmr.hpp:303-368): every block of the update pass derives the same values
__device__ inline void qmr_next(const QmrSmooth& Q, double beta, double rho_new,
                                double xi_new, QmrSmooth* N, double* coeff) {
  const double theta_new = rho_new / (Q.gamma * fabs(beta));
  const double gamma_new = 1.0 / sqrt(1.0 + theta_new * theta_new);
  N->tau = Q.tau * theta_new * gamma_new;
  N->eta = -Q.eta * (Q.rho / beta) * (gamma_new * gamma_new) / (Q.gamma * Q.gamma);
  *coeff = (Q.theta * gamma_new) * (Q.theta * gamma_new);
  N->theta = theta_new;
  N->gamma = gamma_new;
  N->rho = rho_new;
  N->xi = xi_new;
}

struct QmrGateNotDone {
  const int* done;
  __device__ bool operator()() const { return !*done; }
};
struct QmrGateStep {
  const QmrState* st;
  int64_t step;
  __device__ bool operator()() const { return st->k4_step == step; }
};

template <typename T>
struct EpiQmrZParts {  // z = M_R(M_L(A p)), partial <q, z> (summed next)
  static constexpr bool kReduce = true;
  const T* __restrict__ dl;
  const T* __restrict__ dr;
  const T* __restrict__ q;
  T* __restrict__ z;
  const int* done;
  __device__ double row(int i, T s) const {
    T zi = dl ? dl[i] * s : s;
    zi = dr ? dr[i] * zi : zi;
    z[i] = zi;
    return (double)q[i] * (double)zi;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
struct EpiQmrWtParts {  // w~ = A^T u - beta w, partial |w~|^2 (summed next)
  static constexpr bool kReduce = true;
  const T* __restrict__ w;
  T* __restrict__ wt;
  const QmrState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    const T v = s - (T)st->beta * w[i];
    wt[i] = v;
    return (double)v * (double)v;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
struct EpiQmrZ {  // z = M_R(M_L(A p)) for diagonal M, partial <q, z>
  static constexpr bool kReduce = true;
  const T* __restrict__ dl;
  const T* __restrict__ dr;
  const T* __restrict__ q;
  T* __restrict__ z;
  QmrState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    T zi = dl ? dl[i] * s : s;
    zi = dr ? dr[i] * zi : zi;
    z[i] = zi;
    return (double)q[i] * (double)zi;
  }
  __device__ void finalize(double t) const {
    st->eps = t;
    st->beta = t / st->delta;
  }
};

template <typename T>
struct EpiQmrStore {  // work = A p (general preconditioners)
  static constexpr bool kReduce = false;
  T* __restrict__ y;
  const int* done;
  __device__ double row(int i, T s) const {
    y[i] = s;
    return 0.0;
  }
  __device__ void finalize(double) const {}
};

// apply_qmr_smoothing + record_residual + check_convergence
// (qmr.hpp:303-368)
__device__ inline void qmr_smooth(QmrState* S, double xi_new) {
  const double rho_new = S->rho_new;
  const double theta_new = rho_new / (S->gamma * fabs(S->beta));
  const double gamma_new = 1.0 / sqrt(1.0 + theta_new * theta_new);
  S->tau = S->tau * theta_new * gamma_new;
  const double eta_new =
      -S->eta * (S->rho / S->beta) * (gamma_new * gamma_new) / (S->gamma * S->gamma);
  const double theta_old = S->theta;
  S->theta = theta_new;
  S->gamma = gamma_new;
  S->eta = eta_new;
  S->coeff = (theta_old * gamma_new) * (theta_old * gamma_new);
  S->iters += 2;
  S->steps += 1;
  const double est = fabs(S->tau);
  S->resnorm = est;
  if (S->collect) S->history[S->steps - 1] = est;
  if (est / S->bnorm < S->tol) {
    S->converged = 1;
    S->done = 1;
  } else if (rho_new == 0.0 || xi_new == 0.0) {
    S->breakdown = 1;
    S->done = 1;
  }
  S->rho = rho_new;
  S->xi = xi_new;
}

template <typename T>
struct EpiQmrWt {  // w~ = A^T u - beta w, partial |w~|^2; finisher: smoothing
  static constexpr bool kReduce = true;
  const T* __restrict__ w;
  T* __restrict__ wt;
  QmrState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    const T v = s - (T)st->beta * w[i];
    wt[i] = v;
    return (double)v * (double)v;
  }
  __device__ void finalize(double t) const { qmr_smooth(st, sqrt(t)); }
};

template <typename T>
void qmr_solve(spk_csr_s* A, spk_csr_s* At, spk_precond_s* ML, spk_precond_s* MR,
               const spk_solver_config* cfg, const T* b, T* x, spk_summary* out,
               double* history) {
  const int64_t n = A->nrows;
  const size_t vb = (size_t)(n > 0 ? n : 1) * sizeof(T);
  Workspace& ws = workspace();
  T* v = (T*)ws.get(kWsQmr + 0, vb);
  T* w = (T*)ws.get(kWsQmr + 1, vb);
  T* vt = (T*)ws.get(kWsQmr + 2, vb);
  T* wt = (T*)ws.get(kWsQmr + 3, vb);
  T* p = (T*)ws.get(kWsQmr + 4, vb);
  T* q = (T*)ws.get(kWsQmr + 5, vb);
  T* d = (T*)ws.get(kWsQmr + 6, vb);
  T* z = (T*)ws.get(kWsQmr + 7, vb);
  T* work = (T*)ws.get(kWsQmr + 8, vb);
  T* u = (T*)ws.get(kWsQmr + 9, vb);
  T* tmp = (T*)ws.get(kWsQmr + 10, vb);
  const int64_t max_steps = (cfg->max_iterations + 1) / 2;
  double* dhist = cfg->collect_residuals
                      ? (double*)ws.get(kWsQmr + 11, sizeof(double) *
                                        (size_t)std::max<int64_t>(max_steps, 1))
                      : nullptr;
  DevState<QmrState> st(kWsQmr + 12);
  QmrState h{};
  h.tol = cfg->tolerance;
  h.collect = cfg->collect_residuals ? 1 : 0;
  h.history = dhist;
  st.upload(h);
  QmrState* S = st.d;
  hipStream_t s = stream();
  const int* done = &S->done;

  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)b[i] * (double)b[i];
  }, [=] __device__(const double* t) { S->bnorm = sqrt(t[0]); });
  h = st.download();
  if (h.bnorm == 0.0) {  // qmr.hpp:176-180
    launch_map(n, [=] __device__(int64_t i) { x[i] = T(0); });
    *out = spk_summary{0.0, 0, 1, SPK_OK, 0};
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }
  const bool diag = is_diagonal(ML) && is_diagonal(MR);
  const T* dl = diag ? diag_of<T>(ML) : nullptr;
  const T* dr = diag ? diag_of<T>(MR) : nullptr;
  // primal M_R(M_L(in)) or adjoint M_L(M_R(in)) for general preconditioners
  auto precond = [&](const T* in, T* outv, bool adjoint) {
    spk_precond_s* first = adjoint ? MR : ML;
    spk_precond_s* second = adjoint ? ML : MR;
    const bool f = !is_identity(first), g = !is_identity(second);
    if (!f && !g) launch_map(n, [=] __device__(int64_t i) { outv[i] = in[i]; });
    else if (f && !g) first->apply(in, outv);
    else if (!f && g) second->apply(in, outv);
    else {
      first->apply(in, tmp);
      second->apply((const T*)tmp, outv);
    }
  };
  // initialize_lanczos (:189-225): r0 = b - A x; v~ = M_R M_L r0,
  // w~ = M_L M_R r0; rho = |v~|, xi = |w~|; v = v~/rho, w = w~/xi
  launch_spmv(*A, (const T*)x, EpiResidual<T>{b, work}, s);
  precond(work, vt, false);
  precond(work, wt, true);
  launch_reduce<2>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)vt[i] * (double)vt[i];
    acc[1] += (double)wt[i] * (double)wt[i];
  }, [=] __device__(const double* t) {
    S->rho = sqrt(t[0]);
    S->xi = sqrt(t[1]);
    S->gamma = 1.0;
    S->eta = -1.0;
    S->theta = 0.0;
    S->eps = 0.0;
    S->tau = S->rho;
    S->q[0] = QmrSmooth{S->rho, S->xi, 1.0, -1.0, 0.0, S->rho};
    if (S->rho == 0.0) {  // x is exact
      S->converged = 1;
      S->done = 1;
    } else if (S->xi == 0.0) {
      S->breakdown = 1;
      S->done = 1;
    }
  });
  h = st.download();
  // One-kernel hand-off (diagonal preconditioners, few SpMV row blocks):
  // per step  K2 (delta from the partials of the previous update pass;
  // f1, f2; p, q, u)  K3 z = M A p (partials <q, z>)  K4 (epsilon, beta;
  // v~ = z - beta v, partials |v~|^2)  K5 w~ = A^T u - beta w (partials
  // |w~|^2)  K6 (rho_new, xi_new and the smoothing, derived by every block
  // from the previous step's state; d, x, v, w; partials <w, v> for the
  // next delta).  The smoothing state alternates between two slots by
  // step parity; K6 is gated on K4's step stamp, as its block 0 may end
  // the solve.
  const bool handoff = diag && spmv_partials_count(*A) <= kProMaxPartials &&
                       spmv_partials_count(*At) <= kProMaxPartials &&
                       pro_grid(n) <= kProMaxPartials;
  double* pq[4] = {nullptr, nullptr, nullptr, nullptr};
  int c_delta = 0;
  if (handoff && !h.done) {
    const size_t pb = sizeof(double) * 2 * (size_t)kProMaxPartials;
    for (int i = 0; i < 4; ++i) pq[i] = (double*)ws.get(kWsQmrParts + i, pb);
  }
  if (!h.done) {
    if (handoff) {
      c_delta = launch_reduce_pro<1, 0, 1>(
          n, pq[0], 0, (const double*)nullptr, 0, QmrGateNotDone{done},
          [=] __device__(const double*, double*, bool) {},
          [=] __device__(int64_t i) { return Pack4<T>{vt[i], wt[i], T(0), T(0)}; },
          [=] __device__(int64_t i, const Pack4<T>& e, double* acc, const double*) {
            const T vi = e.a * (T)(1.0 / S->rho);
            const T wi = e.b * (T)(1.0 / S->xi);
            v[i] = vi;
            w[i] = wi;
            d[i] = T(0);
            acc[0] += (double)wi * (double)vi;
          }, pq[0]);
    } else {
      launch_map(n, [=] __device__(int64_t i) {
        v[i] = vt[i] * (T)(1.0 / S->rho);
        w[i] = wt[i] * (T)(1.0 / S->xi);
        d[i] = T(0);
      });
    }
    auto issue_handoff = [&](int64_t k) {
      const bool first = k == 0;
      const int par = (int)(k & 1);
      const int64_t step = k + 1;
      // K2: delta (:237-242), f1, f2 (:244-249), p, q, u (:263-281)
      launch_reduce_pro<1, 0, 0>(
          n, pq[0], c_delta, (const double*)nullptr, 0, QmrGateNotDone{done},
          [=] __device__(const double* t, double* sc, bool b0) {
            const double dl0 = t[0];
            sc[2] = dl0 == 0.0 ? 1.0 : 0.0;
            sc[0] = first ? 0.0 : (S->q[par].xi * dl0) / S->eps;
            sc[1] = first ? 0.0 : (S->q[par].rho * dl0) / S->eps;
            if (b0) {
              if (dl0 == 0.0) {
                S->breakdown = 1;
                S->done = 1;
              } else {
                S->delta = dl0;
              }
            }
          },
          [=] __device__(int64_t i) { return Pack4<T>{v[i], w[i], p[i], q[i]}; },
          [=] __device__(int64_t i, const Pack4<T>& e, double*, const double* sc) {
            if (sc[2] != 0.0) return;
            T pi, qi;
            if (first) {
              pi = e.a;
              qi = e.b;
            } else {
              pi = e.a - (T)sc[0] * e.c;
              qi = e.b - (T)sc[1] * e.d;
            }
            p[i] = pi;
            q[i] = qi;
            T ui = dr ? dr[i] * qi : qi;
            u[i] = dl ? dl[i] * ui : ui;
          }, (double*)nullptr);
      // K3: z = M A p, <q, z> (:283-297)
      const int c_eps = launch_spmv_partials(*A, (const T*)p,
                                             EpiQmrZParts<T>{dl, dr, q, z, done}, pq[1], s);
      // K4: epsilon, beta (:298-300); v~ = z - beta v (:301-306)
      const int c_rho = launch_reduce_pro<1, 0, 1>(
          n, pq[1], c_eps, (const double*)nullptr, 0, QmrGateNotDone{done},
          [=] __device__(const double* t, double* sc, bool b0) {
            const double beta = t[0] / S->delta;
            sc[0] = beta;
            if (b0) {
              S->eps = t[0];
              S->beta = beta;
              S->k4_step = step;
            }
          },
          [=] __device__(int64_t i) { return Pack4<T>{z[i], v[i], T(0), T(0)}; },
          [=] __device__(int64_t i, const Pack4<T>& e, double* acc, const double* sc) {
            const T vi = e.a - (T)sc[0] * e.b;
            vt[i] = vi;
            acc[0] += (double)vi * (double)vi;
          }, pq[2]);
      // K5: w~ = A^T u - beta w (:307-320)
      const int c_xi = launch_spmv_partials(*At, (const T*)u,
                                            EpiQmrWtParts<T>{w, wt, S, done}, pq[3], s);
      // K6: smoothing (:303-368), d, x (:330-348), v, w (:370-382), <w, v>
      c_delta = launch_reduce_pro<1, 1, 1>(
          n, pq[2], c_rho, pq[3], c_xi, QmrGateStep{S, step},
          [=] __device__(const double* t, double* sc, bool b0) {
            const double rho_new = sqrt(t[0]), xi_new = sqrt(t[1]);
            QmrSmooth nx{};
            double coeff = 0.0;
            qmr_next(S->q[par], S->beta, rho_new, xi_new, &nx, &coeff);
            sc[0] = nx.eta;
            sc[1] = coeff;
            sc[2] = rho_new;
            sc[3] = xi_new;
            if (b0) {
              S->q[par ^ 1] = nx;
              S->iters += 2;
              S->steps += 1;
              const double est = fabs(nx.tau);
              S->resnorm = est;
              if (S->collect) S->history[S->steps - 1] = est;
              if (est / S->bnorm < S->tol) {
                S->converged = 1;
                S->done = 1;
              } else if (rho_new == 0.0 || xi_new == 0.0) {
                S->breakdown = 1;
                S->done = 1;
              }
            }
          },
          [=] __device__(int64_t i) {
            return Pack8<T>{p[i], d[i], x[i], vt[i], wt[i], T(0), T(0), T(0)};
          },
          [=] __device__(int64_t i, const Pack8<T>& e, double* acc, const double* sc) {
            const T di = (T)sc[0] * e.a + (T)sc[1] * e.b;
            d[i] = di;
            x[i] = e.c + di;
            const T vi = e.d / (T)sc[2];
            const T wi = e.e / (T)sc[3];
            v[i] = vi;
            w[i] = wi;
            acc[0] += (double)wi * (double)vi;
          }, pq[0]);
    };
    auto issue = [&](int64_t k) {
      if (handoff) {
        issue_handoff(k);
        return;
      }
      const bool first = k == 0;
      // qmr_step (:237-261): delta = <w, v>
      launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
        acc[0] += (double)w[i] * (double)v[i];
      }, [=] __device__(const double* t) {
        const double dl0 = t[0];
        if (dl0 == 0.0) {
          S->breakdown = 1;
          S->done = 1;
          return;
        }
        S->delta = dl0;
        if (!first) {
          S->f1 = (S->xi * dl0) / S->eps;
          S->f2 = (S->rho * dl0) / S->eps;
        }
      }, done);
      // update_search_directions (:263-281)
      launch_map(n, [=] __device__(int64_t i) {
        T pi, qi;
        if (first) {
          pi = v[i];
          qi = w[i];
        } else {
          pi = v[i] - (T)S->f1 * p[i];
          qi = w[i] - (T)S->f2 * q[i];
        }
        p[i] = pi;
        q[i] = qi;
        if (diag) {  // adjoint chain M_L(M_R(q))
          T ui = dr ? dr[i] * qi : qi;
          u[i] = dl ? dl[i] * ui : ui;
        }
      }, done);
      // compute_lanczos_step (:283-306)
      if (diag) {
        launch_spmv(*A, (const T*)p, EpiQmrZ<T>{dl, dr, q, z, S, done}, s);
      } else {
        launch_spmv(*A, (const T*)p, EpiQmrStore<T>{work, done}, s);
        precond(work, z, false);
        launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
          acc[0] += (double)q[i] * (double)z[i];
        }, [=] __device__(const double* t) {
          S->eps = t[0];
          S->beta = t[0] / S->delta;
        }, done);
        precond(q, u, true);
      }
      launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
        const T vi = z[i] - (T)S->beta * v[i];
        vt[i] = vi;
        acc[0] += (double)vi * (double)vi;
      }, [=] __device__(const double* t) { S->rho_new = sqrt(t[0]); }, done);
      launch_spmv(*At, (const T*)u, EpiQmrWt<T>{w, wt, S, done}, s);
      // update_solution (:330-348) + advance_lanczos (:370-382)
      const int64_t step = k + 1;
      launch_map(n, [=] __device__(int64_t i) {
        if (S->steps != step) return;
        const T di = (T)S->eta * p[i] + (T)S->coeff * d[i];
        d[i] = di;
        x[i] += di;
        v[i] = vt[i] / (T)S->rho;
        w[i] = wt[i] / (T)S->xi;
      });
    };
    auto is_done = [&]() { return st.download().done != 0; };
    chunked_loop(max_steps, issue, is_done, 2, 32);
  }
  h = st.download();
  out->residual_norm = h.resnorm;
  out->computed_iterations = h.iters;
  out->converged = h.converged;
  out->status = h.breakdown ? SPK_BREAKDOWN : SPK_OK;
  out->history_len = h.collect ? h.steps : 0;
  if (history && h.collect && h.steps > 0) {
    SPK_HIP_TRY(hipMemcpyAsync(history, dhist, h.steps * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  }
  SPK_HIP_TRY(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_qmr_solve(spk_csr A, spk_csr At, spk_precond M_left,
                                    spk_precond M_right,
                                    const spk_solver_config* cfg, const void* d_b,
                                    void* d_x, spk_summary* summary,
                                    double* history) {
  return guarded([&] {
    TraceRange trace_("spk_qmr_solve");
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(At && summary, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(At->nrows == A->ncols && At->ncols == A->nrows && At->nnz == A->nnz,
                SPK_INVALID_ARGUMENT, "At must be the transpose of A");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      qmr_solve<float>(A, At, M_left, M_right, cfg, (const float*)d_b,
                       (float*)d_x, summary, history);
    else
      qmr_solve<double>(A, At, M_left, M_right, cfg, (const double*)d_b,
                        (double*)d_x, summary, history);
    check_preconds(M_left, M_right);
  });
}
