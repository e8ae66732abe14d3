// spk_minres.hip -- device MINRES (minres.hpp:80-395): Lanczos with the
// combined preconditioner z = M_R(M_L(v)), a QR factorisation by Givens
// rotations kept on the device, and the d-vector solution update.  Per step:
//   K1  w = A z, <z, w> in the epilogue;  finisher: alpha
//   K2  w -= alpha v + beta v_prev;  u = M_R(M_L(w)) (fused for diagonal
//       preconditioners, written over v_prev); <w, u>;  finisher:
//       beta_new, the two old rotations and the new one, phi, phi_bar,
//       the residual estimate |phi_bar| and the convergence test
//   K3  d_new = (z - delta d - epsilon d_prev) / rho;  x += phi d_new;
//       z = u / beta_new;  v_new = w / beta_new (gated on the step count:
//       it must also run in the converging step)
// d/d_prev and v/v_prev trade places by renaming buffers on the host.
// rho == 0 (the reference then updates with the previous step's rho,
// :293-296) stops with SPK_BREAKDOWN.
#include "spk_solver_common.hpp"

namespace spk {
namespace {

// the rotation state one step hands to the next
struct MinresGivens {
  double beta, cs_prev, sn_prev, cs, sn, phi_bar;
};

struct MinresState {
  double bnorm, tol, resnorm, beta, alpha, cs_prev, sn_prev, cs, sn;
  double phi, phi_bar, delta, eps, rho, inv_bn;
  int64_t iters;
  int32_t done, converged, breakdown, collect;
  double* history;
  // one-kernel hand-off: the rotation state by step parity, and the step
  // whose w pass ran (gates the update pass)
  MinresGivens g[2];
  int64_t k2_step;
};

// minres_givens' arithmetic as a pure function of the previous step's
// rotation state, alpha and <w, u> (minres.hpp:278-360): every block of the
// update pass derives the same values
struct MinresRot {
  double delta, eps, rho, phi, inv_bn, est, beta_new;
  bool breakdown;
};
__device__ inline MinresRot minres_rot(const MinresGivens& G, double alpha, double wu,
                                       MinresGivens* next) {
  MinresRot r{};
  r.beta_new = sqrt(wu);
  r.eps = G.sn_prev * G.beta;
  const double delta_1 = G.cs_prev * G.beta;
  r.delta = G.cs * delta_1 + G.sn * alpha;
  const double rho_bar = -G.sn * delta_1 + G.cs * alpha;
  r.rho = sqrt(rho_bar * rho_bar + r.beta_new * r.beta_new);
  r.breakdown = r.rho == 0.0;
  if (r.breakdown) return r;
  const double cs_new = rho_bar / r.rho, sn_new = r.beta_new / r.rho;
  r.phi = cs_new * G.phi_bar;
  next->phi_bar = -sn_new * G.phi_bar;
  r.est = fabs(next->phi_bar);
  r.inv_bn = r.beta_new != 0.0 ? 1.0 / r.beta_new : 0.0;
  next->cs_prev = G.cs;
  next->sn_prev = G.sn;
  next->cs = cs_new;
  next->sn = sn_new;
  next->beta = r.beta_new;
  return r;
}

struct MinGateNotDone {
  const int* done;
  __device__ bool operator()() const { return !*done; }
};
struct MinGateStep {
  const MinresState* st;
  int64_t step;
  __device__ bool operator()() const { return st->k2_step == step; }
};

template <typename T>
struct EpiMinresWParts {  // w = A z, partial <z, w> (summed by the next kernel)
  static constexpr bool kReduce = true;
  const T* __restrict__ z;
  T* __restrict__ w;
  const int* done;
  __device__ double row(int i, T s) const {
    w[i] = s;
    return (double)z[i] * (double)s;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
struct EpiMinresW {  // w = A z, partial <z, w>; finisher: alpha
  static constexpr bool kReduce = true;
  const T* __restrict__ z;
  T* __restrict__ w;
  MinresState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    w[i] = s;
    return (double)z[i] * (double)s;
  }
  __device__ void finalize(double t) const { st->alpha = t; }
};

// apply_givens_rotations + check_convergence + rotate_state's scalars
// (minres.hpp:278-360)
__device__ inline void minres_givens(MinresState* S, double wu) {
  const double beta_new = sqrt(wu);
  const double eps = S->sn_prev * S->beta;
  const double delta_1 = S->cs_prev * S->beta;
  const double delta = S->cs * delta_1 + S->sn * S->alpha;
  const double rho_bar = -S->sn * delta_1 + S->cs * S->alpha;
  const double rho = sqrt(rho_bar * rho_bar + beta_new * beta_new);
  if (rho == 0.0) {
    S->breakdown = 1;
    S->done = 1;
    return;
  }
  const double cs_new = rho_bar / rho, sn_new = beta_new / rho;
  S->phi = cs_new * S->phi_bar;
  S->phi_bar = -sn_new * S->phi_bar;
  S->rho = rho;
  S->delta = delta;
  S->eps = eps;
  const double est = fabs(S->phi_bar);
  S->resnorm = est;
  if (S->collect) S->history[S->iters] = est;
  S->iters += 1;
  if (est / S->bnorm < S->tol) {
    S->converged = 1;
    S->done = 1;
  } else if (beta_new == 0.0) {
    S->done = 1;  // invariant Krylov space: phi_bar is zero up to rounding
  }
  S->inv_bn = beta_new != 0.0 ? 1.0 / beta_new : 0.0;
  S->cs_prev = S->cs;
  S->sn_prev = S->sn;
  S->cs = cs_new;
  S->sn = sn_new;
  S->beta = beta_new;
}

template <typename T>
void minres_solve(spk_csr_s* A, spk_precond_s* ML, spk_precond_s* MR,
                  const spk_solver_config* cfg, const T* b, T* x,
                  spk_summary* out, double* history) {
  const int64_t n = A->nrows;
  const size_t vb = (size_t)(n > 0 ? n : 1) * sizeof(T);
  Workspace& ws = workspace();
  T* vbuf[2] = {(T*)ws.get(kWsMinres + 0, vb), (T*)ws.get(kWsMinres + 1, vb)};
  T* dbuf[2] = {(T*)ws.get(kWsMinres + 2, vb), (T*)ws.get(kWsMinres + 3, vb)};
  T* z = (T*)ws.get(kWsMinres + 4, vb);
  T* w = (T*)ws.get(kWsMinres + 5, vb);
  T* tmp = (T*)ws.get(kWsMinres + 6, vb);
  double* dhist = cfg->collect_residuals
                      ? (double*)ws.get(kWsMinres + 7, sizeof(double) *
                                        (size_t)std::max<int64_t>(cfg->max_iterations, 1))
                      : nullptr;
  DevState<MinresState> st(kWsMinres + 8);
  MinresState h{};
  h.tol = cfg->tolerance;
  h.collect = cfg->collect_residuals ? 1 : 0;
  h.history = dhist;
  st.upload(h);
  MinresState* S = st.d;
  hipStream_t s = stream();
  const int* done = &S->done;

  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)b[i] * (double)b[i];
  }, [=] __device__(const double* t) { S->bnorm = sqrt(t[0]); });
  h = st.download();
  if (h.bnorm == 0.0) {  // minres.hpp:170-174
    launch_map(n, [=] __device__(int64_t i) { x[i] = T(0); });
    *out = spk_summary{0.0, 0, 1, SPK_OK, 0};
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }
  const bool diag = is_diagonal(ML) && is_diagonal(MR);
  const T* dl = diag ? diag_of<T>(ML) : nullptr;
  const T* dr = diag ? diag_of<T>(MR) : nullptr;
  // out = M_R(M_L(in)) for general preconditioners
  auto precond = [&](const T* in, T* outv) {
    const bool f = !is_identity(ML), g = !is_identity(MR);
    if (!f && !g) launch_map(n, [=] __device__(int64_t i) { outv[i] = in[i]; });
    else if (f && !g) ML->apply(in, outv);
    else if (!f && g) MR->apply(in, outv);
    else {
      ML->apply(in, tmp);
      MR->apply((const T*)tmp, outv);
    }
  };
  T* v = vbuf[0];
  T* vp = vbuf[1];
  T* d = dbuf[0];
  T* dp = dbuf[1];
  // compute_initial_residual (:180-213): v = b - A x; z = M v;
  // beta = sqrt(<v, z>); v /= beta, z /= beta
  launch_spmv(*A, (const T*)x, EpiResidual<T>{b, v}, s);
  precond(v, z);
  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)v[i] * (double)z[i];
  }, [=] __device__(const double* t) {
    S->beta = sqrt(t[0]);
    S->phi_bar = S->beta;
    S->cs_prev = 1.0;
    S->sn_prev = 0.0;
    S->cs = 1.0;
    S->sn = 0.0;
    S->g[0] = MinresGivens{S->beta, 1.0, 0.0, 1.0, 0.0, S->beta};
    if (S->beta == 0.0) {
      S->converged = 1;
      S->done = 1;
    }
  });
  h = st.download();
  if (!h.done) {
    launch_map(n, [=] __device__(int64_t i) {
      const T inv = (T)(1.0 / S->beta);
      v[i] = v[i] * inv;
      z[i] = z[i] * inv;
      vp[i] = T(0);
      d[i] = T(0);
      dp[i] = T(0);
    });
    // One-kernel hand-off (diagonal preconditioners, few SpMV row blocks):
    // w = A z (partials <z, w>); the w pass sums them (alpha) and writes
    // partials <w, u>; the update pass sums those, and every block derives
    // the rotations from the previous step's state (block 0 records the new
    // state in the other parity slot).
    const bool handoff = diag && spmv_partials_count(*A) <= kProMaxPartials &&
                         pro_grid(n) <= kProMaxPartials;
    double* pq[2] = {nullptr, nullptr};
    if (handoff) {
      const size_t pb = sizeof(double) * 2 * (size_t)kProMaxPartials;
      for (int i = 0; i < 2; ++i) pq[i] = (double*)ws.get(kWsMinresParts + i, pb);
    }
    auto issue_handoff = [&](int64_t k) {
      const int par = (int)(k & 1);
      const int64_t step = k + 1;
      const int c1 = launch_spmv_partials(*A, (const T*)z,
                                          EpiMinresWParts<T>{z, w, done}, pq[0], s);
      T* u = vp;
      const T* vc = v;
      const int c2 = launch_reduce_pro<1, 0, 1>(
          n, pq[0], c1, (const double*)nullptr, 0, MinGateNotDone{done},
          [=] __device__(const double* t, double* sc, bool b0) {
            sc[0] = t[0];
            sc[1] = S->g[par].beta;
            if (b0) {
              S->alpha = t[0];
              S->k2_step = step;
            }
          },
          [=] __device__(int64_t i) {
            return Pack8<T>{w[i], vc[i], u[i], dl ? dl[i] : T(1), dr ? dr[i] : T(1),
                            T(0), T(0), T(0)};
          },
          [=] __device__(int64_t i, const Pack8<T>& e, double* acc, const double* sc) {
            const T wi = e.a - ((T)sc[0] * e.b + (T)sc[1] * e.c);
            w[i] = wi;
            T ui = dl ? e.d * wi : wi;
            ui = dr ? e.e * ui : ui;
            u[i] = ui;
            acc[0] += (double)wi * (double)ui;
          }, pq[1]);
      T* dc = d;
      T* dn = dp;
      T* vn = vp;
      launch_reduce_pro<1, 0, 0>(
          n, pq[1], c2, (const double*)nullptr, 0, MinGateStep{S, step},
          [=] __device__(const double* t, double* sc, bool b0) {
            MinresGivens next{};
            const MinresRot r = minres_rot(S->g[par], S->alpha, t[0], &next);
            sc[0] = r.delta;
            sc[1] = r.eps;
            sc[2] = r.rho;
            sc[3] = r.phi;
            sc[4] = r.inv_bn;
            sc[5] = r.breakdown ? 0.0 : 1.0;
            if (b0) {
              if (r.breakdown) {
                S->breakdown = 1;
                S->done = 1;
              } else {
                S->g[par ^ 1] = next;
                S->resnorm = r.est;
                if (S->collect) S->history[S->iters] = r.est;
                S->iters += 1;
                if (r.est / S->bnorm < S->tol) {
                  S->converged = 1;
                  S->done = 1;
                } else if (r.beta_new == 0.0) {
                  S->done = 1;
                }
              }
            }
          },
          [=] __device__(int64_t i) {
            return Pack8<T>{z[i], dc[i], dn[i], vn[i], w[i], x[i], T(0), T(0)};
          },
          [=] __device__(int64_t i, const Pack8<T>& e, double*, const double* sc) {
            if (sc[5] == 0.0) return;
            const T dnew = (e.a - (T)sc[0] * e.b - (T)sc[1] * e.c) / (T)sc[2];
            dn[i] = dnew;
            x[i] = e.f + (T)sc[3] * dnew;
            const T inv = (T)sc[4];
            z[i] = e.d * inv;
            vn[i] = e.e * inv;
          }, (double*)nullptr);
      std::swap(d, dp);
      std::swap(v, vp);
    };
    auto issue = [&](int64_t k) {
      if (handoff) {
        issue_handoff(k);
        return;
      }
      // lanczos_step (:227-237)
      launch_spmv(*A, (const T*)z, EpiMinresW<T>{z, w, S, done}, s);
      // precondition_lanczos (:239-250) + the rotations (:252-275)
      T* u = vp;
      const T* vc = v;
      if (diag) {
        launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
          const T wi = w[i] - ((T)S->alpha * vc[i] + (T)S->beta * u[i]);
          w[i] = wi;
          T ui = dl ? dl[i] * wi : wi;
          ui = dr ? dr[i] * ui : ui;
          u[i] = ui;
          acc[0] += (double)wi * (double)ui;
        }, [=] __device__(const double* t) { minres_givens(S, t[0]); }, done);
      } else {
        launch_map(n, [=] __device__(int64_t i) {
          w[i] = w[i] - ((T)S->alpha * vc[i] + (T)S->beta * u[i]);
        }, done);
        precond(w, u);
        launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
          acc[0] += (double)w[i] * (double)u[i];
        }, [=] __device__(const double* t) { minres_givens(S, t[0]); }, done);
      }
      // update_solution (:277-286) + rotate_state's vectors (:298-310)
      const int64_t step = k + 1;
      T* dc = d;
      T* dn = dp;
      T* vn = vp;
      launch_map(n, [=] __device__(int64_t i) {
        if (S->iters != step) return;
        const T dnew = (z[i] - (T)S->delta * dc[i] - (T)S->eps * dn[i]) / (T)S->rho;
        dn[i] = dnew;
        x[i] += (T)S->phi * dnew;
        const T inv = (T)S->inv_bn;
        z[i] = vn[i] * inv;
        vn[i] = w[i] * inv;
      });
      std::swap(d, dp);
      std::swap(v, vp);
    };
    auto is_done = [&]() { return st.download().done != 0; };
    chunked_loop(cfg->max_iterations, issue, is_done);
  }
  h = st.download();
  out->residual_norm = h.resnorm;
  out->computed_iterations = h.iters;
  out->converged = h.converged;
  out->status = h.breakdown ? SPK_BREAKDOWN : SPK_OK;
  out->history_len = h.collect ? h.iters : 0;
  if (history && h.collect && h.iters > 0) {
    SPK_HIP_TRY(hipMemcpyAsync(history, dhist, h.iters * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  }
  SPK_HIP_TRY(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_minres_solve(spk_csr A, spk_precond M_left,
                                       spk_precond M_right,
                                       const spk_solver_config* cfg,
                                       const void* d_b, void* d_x,
                                       spk_summary* summary, double* history) {
  return guarded([&] {
    TraceRange trace_("spk_minres_solve");
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(summary, SPK_INVALID_ARGUMENT, "null summary");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      minres_solve<float>(A, M_left, M_right, cfg, (const float*)d_b,
                          (float*)d_x, summary, history);
    else
      minres_solve<double>(A, M_left, M_right, cfg, (const double*)d_b,
                           (double*)d_x, summary, history);
    check_preconds(M_left, M_right);
  });
}
