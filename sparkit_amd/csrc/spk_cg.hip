// spk_cg.hip -- device conjugate gradient with the reference's unified
// left/right preconditioned recurrence (conjugate_gradient.hpp:122-312):
//
//   every restart_iterations steps: r = b - A x; z = M_L r; rz = <r,z>; p = z
//   p^ = M_R p; q = A p^; alpha = rz / <p,q>
//   x += alpha p^; r -= alpha q; |r| / |b| < tol -> stop
//   z = M_L r; beta = <r,z>_new / rz; p = z + beta p
//
// Scalars (rz, alpha, beta, |r|, the iteration count and the convergence
// flag) live on the device and are updated by the one-block finisher that
// follows each fused reduction kernel; the host only queues kernels and
// polls the flag once per chunk.  With M_L = Jacobi and M_R = identity
// (Jacobi-PCG) one iteration is three kernels plus two finishers:
//   K1  q = A p, partial <p,q>;  finisher: alpha = rz/<p,q>
//   K2  r -= alpha q; <r,r>, <r, D^-1 r>;  finisher: |r|, convergence flag,
//       beta, rz
//   K3  x += alpha p; p = D^-1 r + beta p
// K3 is the one pass that reads the old p, so the x update rides there.  It
// must also run in the step whose K2 raised the convergence flag, so it is
// gated on the step count (it runs iff K2 of the same step ran) instead of
// the flag.
// HBM traffic per iteration: B_spmv + 10 s n bytes (K2 reads r,q,D^-1 and
// writes r; K3 reads x,r,D^-1,p and writes x,p).
// With an explicit approximate inverse M as left preconditioner the same
// shape holds with one more SpMV: K2 updates r and |r|; z = M r runs as an
// SpMV whose epilogue stores z and sums <r,z> (finisher: beta, rz); the
// last pass does x += alpha p; p = z + beta p.  (Forming r - alpha q inside
// that SpMV's gather instead of K2 was measured slower: two loads per
// gathered column cost more than the pass and launch they save.)
// When the SpMV has at most kProMaxPartials row blocks (and M_R is the
// identity, M_L Jacobi or an explicit inverse) the finishers go: K2 sums
// K1's partials itself and derives alpha, the last pass sums K2's (and the
// preconditioner SpMV's) and derives |r|, the convergence test and beta.
// Every block sums the same partials in the same order, so all derive the
// same scalars; block 0 records them.  rz alternates between two slots by
// step parity, as the kernel that reads rz_k writes rz_{k+1}.  Jacobi-PCG
// with more SpMV row blocks than that keeps K1's finisher (alpha from the
// parity slot) and hands off from K2 on.
#include "spk_solver_common.hpp"

namespace spk {
namespace {

struct CgState {
  double bnorm, rz, alpha, beta, resnorm, tol;
  int64_t iters;
  int32_t done, converged, collect, pad;
  double* history;
  // one-kernel hand-off path: rz by step parity (a kernel reads one and
  // writes the other), the step whose residual update ran
  double rzb[2];
  int64_t k2_step;
};

// launch gates of the one-kernel hand-off path
struct GateNotDone {
  const int* done;
  __device__ bool operator()() const { return !*done; }
};
struct GateStep {  // runs iff the residual update of this step ran
  const CgState* st;
  int64_t step;
  __device__ bool operator()() const { return st->k2_step == step; }
};

template <typename T>
struct EpiCgDot {  // q = A p^, partial <p, q>
  static constexpr bool kReduce = true;
  const T* __restrict__ p;
  T* __restrict__ q;
  CgState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    q[i] = s;
    return (double)p[i] * (double)s;
  }
  __device__ void finalize(double pq) const { st->alpha = st->rz / pq; }
};

template <typename T>
struct EpiCgDotRzb {  // q = A p^, alpha = <r,z> / <p,q> from the parity buffer
  static constexpr bool kReduce = true;
  const T* __restrict__ p;
  T* __restrict__ q;
  CgState* st;
  const int* done;
  int par;
  __device__ double row(int i, T s) const {
    q[i] = s;
    return (double)p[i] * (double)s;
  }
  __device__ void finalize(double pq) const { st->alpha = st->rzb[par] / pq; }
};

template <typename T>
struct EpiCgRestartJacobi {  // r = b - A x; z = D^-1 r; p = z; <r, z>
  static constexpr bool kReduce = true;
  const T* __restrict__ b;
  const T* __restrict__ invd;
  T* __restrict__ r;
  T* __restrict__ p;
  CgState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    T ri = b[i] - s;
    T zi = invd[i] * ri;
    r[i] = ri;
    p[i] = zi;
    return (double)ri * (double)zi;
  }
  __device__ void finalize(double rz) const {
    st->rz = rz;
    st->rzb[par] = rz;
  }
  int par;  // step parity (the one-kernel hand-off path)
};

template <typename T>
struct EpiCgPrecDot {  // z = M r (explicit approximate inverse), <r, z>
  static constexpr bool kReduce = true;
  const T* __restrict__ r;
  T* __restrict__ z;
  CgState* st;
  const int* done;
  __device__ double row(int i, T s) const {
    z[i] = s;
    return (double)r[i] * (double)s;
  }
  __device__ void finalize(double rz) const {
    st->beta = rz / st->rz;
    st->rz = rz;
  }
};

template <typename T>
struct EpiCgResidual {  // r = b - A x (generic restart)
  static constexpr bool kReduce = false;
  const T* __restrict__ b;
  T* __restrict__ r;
  const int* done;
  __device__ double row(int i, T s) const {
    r[i] = b[i] - s;
    return 0.0;
  }
  __device__ void finalize(double) const {}
};

// is_converged(), conjugate_gradient.hpp:261-276
__device__ inline void cg_check(CgState* st, double rr) {
  double rn = sqrt(rr);
  st->resnorm = rn;
  if (st->collect) st->history[st->iters] = rn;
  st->iters += 1;
  if (rn / st->bnorm < st->tol) {
    st->converged = 1;
    st->done = 1;
  }
}

template <typename T>
void cg_solve(spk_csr_s* A, spk_precond_s* ML, spk_precond_s* MR,
              const spk_solver_config* cfg, const T* b, T* x,
              spk_summary* out, double* history) {
  SPK_REQUIRE(cfg->restart > 0, SPK_INVALID_ARGUMENT,
              "restart_iterations must be > 0 (the reference divides by it)");
  // an inner-solve preconditioner runs a solve of its own (an inner PCG
  // this very solver and its work vectors): it belongs in FGMRES's flexible
  // slot, not in CG's
  SPK_REQUIRE((!ML || !ML->runs_solve()) && (!MR || !MR->runs_solve()),
              SPK_INVALID_ARGUMENT,
              "conjugate_gradient: an inner-PCG or inner-solve preconditioner "
              "is not a fixed operator (use it in fgmres)");
  const int64_t n = A->nrows;
  Workspace& ws = workspace();
  const size_t vb = (size_t)(n > 0 ? n : 1) * sizeof(T);
  T* r = (T*)ws.get(kWsCg + 0, vb);
  T* z = (T*)ws.get(kWsCg + 1, vb);
  T* p = (T*)ws.get(kWsCg + 2, vb);
  T* q = (T*)ws.get(kWsCg + 3, vb);
  const bool right_id = is_identity(MR);
  T* hatp = right_id ? p : (T*)ws.get(kWsCg + 4, vb);
  double* dhist = cfg->collect_residuals
                      ? (double*)ws.get(kWsCg + 5, sizeof(double) *
                                        (size_t)std::max<int64_t>(cfg->max_iterations, 1))
                      : nullptr;
  DevState<CgState> st(kWsCg + 6);
  CgState h{};
  h.tol = cfg->tolerance;
  h.collect = cfg->collect_residuals ? 1 : 0;
  h.history = dhist;
  st.upload(h);
  CgState* S = st.d;
  hipStream_t s = stream();

  // bnorm = sqrt(<b,b>) (:194); zero rhs -> x = 0, converged (:195-199)
  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)b[i] * (double)b[i];
  }, [=] __device__(const double* t) { S->bnorm = sqrt(t[0]); });
  h = st.download();
  if (h.bnorm == 0.0) {
    launch_map(n, [=] __device__(int64_t i) { x[i] = T(0); });
    out->converged = 1;
    out->residual_norm = 0.0;
    out->computed_iterations = 0;
    out->status = SPK_OK;
    out->history_len = 0;
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }

  const bool jac = is_jacobi(ML);
  const T* invd = jac ? ML->inv_diag<T>() : nullptr;
  const bool spmv_left = ML && ML->kind == spk_precond_s::kSpmv;
  const int* done = &S->done;
  // the one-kernel hand-off needs few partials: every block of the next
  // kernel reads all of them
  const bool handoff = right_id && (jac || spmv_left) &&
                       spmv_partials_count(*A) <= kProMaxPartials &&
                       (!spmv_left || spmv_partials_count(*ML->M) <= kProMaxPartials) &&
                       pro_grid(n) <= kProMaxPartials;
  // Jacobi-PCG whose SpMV has too many row blocks for a prologue: the SpMV
  // keeps its finisher (alpha), the r update hands its partials to the
  // x / p pass
  const bool mixed = !handoff && right_id && jac && pro_grid(n) <= kProMaxPartials;
  double* part1 = nullptr;
  double* part2 = nullptr;
  double* part3 = nullptr;
  if (handoff || mixed) {
    const size_t pb = sizeof(double) * 2 * (size_t)kProMaxPartials;
    part1 = (double*)ws.get(kWsCg + 7, pb);
    part2 = (double*)ws.get(kWsCg + 8, pb);
    part3 = (double*)ws.get(kWsCg + 9, pb);
  }

  auto issue = [&](int64_t k) {
    if (k % cfg->restart == 0) {  // reorthogonalize(), :216-230
      if (jac) {
        launch_spmv(*A, (const T*)x,
                    EpiCgRestartJacobi<T>{b, invd, r, p, S, done, (int)(k & 1)}, s);
      } else {
        launch_spmv(*A, (const T*)x, EpiCgResidual<T>{b, r, done}, s);
        apply_precond(ML, n, (const T*)r, z);
        const int par = (int)(k & 1);
        launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
          T zi = z[i];
          p[i] = zi;
          acc[0] += (double)r[i] * (double)zi;
        }, [=] __device__(const double* t) {
          S->rz = t[0];
          S->rzb[par] = t[0];
        }, done);
      }
    }
    if (handoff || mixed) {
      // One-kernel hand-off: no finisher kernels.  K1 q = A p (partials
      // <p,q>); K2 sums them itself -> alpha, r -= alpha q, partials;
      // [K3 z = M r, partials <r,z>]; last pass sums the partials ->
      // |r|, convergence, beta; x += alpha p; p = z + beta p.  Mixed: K1
      // ends with a finisher that stores alpha, which K2 reads.
      const int par = (int)(k & 1);
      const int64_t step = k + 1;
      int c1 = 0;
      if (handoff)
        c1 = launch_spmv_partials(*A, (const T*)p, EpiCgDot<T>{p, q, S, done}, part1, s);
      else
        launch_spmv(*A, (const T*)p, EpiCgDotRzb<T>{p, q, S, done, par}, s);
      const bool ho = handoff;
      auto pro_alpha = [=] __device__(const double* t, double* sc, bool b0) {
        const double alpha = ho ? S->rzb[par] / t[0] : S->alpha;
        sc[0] = alpha;
        if (b0) {
          if (ho) S->alpha = alpha;
          S->k2_step = step;
        }
      };
      // compute_new_solution's |r| and the new beta (:241-292), once per
      // block; block 0 records them (nothing this kernel reads)
      auto pro_beta = [=] __device__(const double* t, double* sc, bool b0) {
        const double rn = sqrt(t[0]);
        const bool conv = rn / S->bnorm < S->tol;
        const double beta = t[1] / S->rzb[par];
        sc[0] = S->alpha;
        sc[1] = beta;
        if (b0) {
          S->resnorm = rn;
          if (S->collect) S->history[S->iters] = rn;
          S->iters += 1;
          if (conv) {
            S->converged = 1;
            S->done = 1;
          } else {
            S->beta = beta;
            S->rzb[par ^ 1] = t[1];
          }
        }
      };
      if (jac) {
        const int g2 = launch_reduce_pro<1, 0, 2>(
            n, part1, c1, (const double*)nullptr, 0, GateNotDone{done}, pro_alpha,
            [=] __device__(int64_t i) { return Pack4<T>{r[i], q[i], invd[i], T(0)}; },
            [=] __device__(int64_t i, const Pack4<T>& e, double* acc, const double* sc) {
              const T a = (T)sc[0];
              T ri = e.a - a * e.b;
              r[i] = ri;
              acc[0] += (double)ri * (double)ri;
              acc[1] += (double)ri * (double)(e.c * ri);
            }, part2);
        launch_reduce_pro<2, 0, 0>(
            n, part2, g2, (const double*)nullptr, 0, GateStep{S, step}, pro_beta,
            [=] __device__(int64_t i) { return Pack4<T>{p[i], x[i], invd[i], r[i]}; },
            [=] __device__(int64_t i, const Pack4<T>& e, double*, const double* sc) {
              x[i] = (T)sc[0] * e.a + e.b;
              p[i] = e.c * e.d + (T)sc[1] * e.a;
            }, (double*)nullptr);
      } else {  // explicit approximate inverse
        const int g2 = launch_reduce_pro<1, 0, 1>(
            n, part1, c1, (const double*)nullptr, 0, GateNotDone{done}, pro_alpha,
            [=] __device__(int64_t i) { return Pack4<T>{r[i], q[i], T(0), T(0)}; },
            [=] __device__(int64_t i, const Pack4<T>& e, double* acc, const double* sc) {
              const T a = (T)sc[0];
              T ri = e.a - a * e.b;
              r[i] = ri;
              acc[0] += (double)ri * (double)ri;
            }, part2);
        const int c3 = launch_spmv_partials(*ML->M, (const T*)r,
                                            EpiCgPrecDot<T>{r, z, S, done}, part3, s);
        launch_reduce_pro<1, 1, 0>(
            n, part2, g2, part3, c3, GateStep{S, step}, pro_beta,
            [=] __device__(int64_t i) { return Pack4<T>{p[i], x[i], z[i], T(0)}; },
            [=] __device__(int64_t i, const Pack4<T>& e, double*, const double* sc) {
              x[i] = (T)sc[0] * e.a + e.b;
              p[i] = e.c + (T)sc[1] * e.a;
            }, (double*)nullptr);
      }
      return;
    }
    // compute_step_size(), :232-239
    if (!right_id) apply_precond(MR, n, (const T*)p, hatp);
    launch_spmv(*A, (const T*)hatp, EpiCgDot<T>{p, q, S, done}, s);
    if (jac && right_id) {
      // compute_new_residual + is_converged + the z, rz_new of
      // update_direction (:248-292), fused
      launch_reduce<2>(n, [=] __device__(int64_t i, double* acc) {
        const T a = (T)S->alpha;
        T ri = r[i] - a * q[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
        acc[1] += (double)ri * (double)(invd[i] * ri);
      }, [=] __device__(const double* t) {
        cg_check(S, t[0]);
        if (!S->done) {
          S->beta = t[1] / S->rz;
          S->rz = t[1];
        }
      }, done);
      // compute_new_solution (:241-246) + the direction update (:278-292);
      // after the last step the new direction is computed and unused
      const int64_t step = k + 1;
      launch_map(n, [=] __device__(int64_t i) {
        if (S->iters != step) return;
        const T pi = p[i];
        x[i] = (T)S->alpha * pi + x[i];
        p[i] = invd[i] * r[i] + (T)S->beta * pi;
      });
    } else if (spmv_left && right_id) {
      // compute_new_residual + is_converged (:248-276)
      launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
        const T a = (T)S->alpha;
        T ri = r[i] - a * q[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
      }, [=] __device__(const double* t) { cg_check(S, t[0]); }, done);
      // z = M r with rz_new = <r, z> in the SpMV epilogue (:278-292)
      launch_spmv(*ML->M, (const T*)r, EpiCgPrecDot<T>{r, z, S, done}, s);
      // compute_new_solution (:241-246) + the direction update, gated on
      // the step count like the Jacobi path's
      const int64_t step = k + 1;
      launch_map(n, [=] __device__(int64_t i) {
        if (S->iters != step) return;
        const T pi = p[i];
        x[i] = (T)S->alpha * pi + x[i];
        p[i] = z[i] + (T)S->beta * pi;
      });
    } else if (jac) {
      // compute_new_solution/residual + is_converged + the z, rz_new of
      // update_direction (:241-292), fused
      launch_reduce<2>(n, [=] __device__(int64_t i, double* acc) {
        const T a = (T)S->alpha;
        x[i] = a * hatp[i] + x[i];
        T ri = r[i] - a * q[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
        acc[1] += (double)ri * (double)(invd[i] * ri);
      }, [=] __device__(const double* t) {
        cg_check(S, t[0]);
        if (!S->done) {
          S->beta = t[1] / S->rz;
          S->rz = t[1];
        }
      }, done);
      launch_map(n, [=] __device__(int64_t i) {
        p[i] = invd[i] * r[i] + (T)S->beta * p[i];
      }, done);
    } else {
      launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
        const T a = (T)S->alpha;
        x[i] = a * hatp[i] + x[i];
        T ri = r[i] - a * q[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
      }, [=] __device__(const double* t) { cg_check(S, t[0]); }, done);
      apply_precond(ML, n, (const T*)r, z);
      launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
        acc[0] += (double)r[i] * (double)z[i];
      }, [=] __device__(const double* t) {
        S->beta = t[0] / S->rz;
        S->rz = t[0];
      }, done);
      launch_map(n, [=] __device__(int64_t i) {
        p[i] = z[i] + (T)S->beta * p[i];
      }, done);
    }
  };
  auto is_done = [&]() { return st.download().done != 0; };
  chunked_loop(cfg->max_iterations, issue, is_done);

  h = st.download();
  out->residual_norm = h.resnorm;
  out->computed_iterations = h.iters;
  out->converged = h.converged;
  out->status = SPK_OK;
  out->history_len = h.collect ? h.iters : 0;
  if (history && h.collect && h.iters > 0) {
    SPK_HIP_TRY(hipMemcpyAsync(history, dhist, h.iters * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
  }
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_cg_solve(spk_csr A, spk_precond M_left,
                                   spk_precond M_right,
                                   const spk_solver_config* cfg,
                                   const void* d_b, void* d_x,
                                   spk_summary* summary, double* history) {
  return guarded([&] {
    TraceRange trace_("spk_cg_solve");
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(summary, SPK_INVALID_ARGUMENT, "null summary");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      cg_solve<float>(A, M_left, M_right, cfg, (const float*)d_b, (float*)d_x,
                      summary, history);
    else
      cg_solve<double>(A, M_left, M_right, cfg, (const double*)d_b,
                       (double*)d_x, summary, history);
    check_preconds(M_left, M_right);
  });
}

namespace spk {
namespace {

// The no-allocation solve may only touch what already exists: a store's
// lazily made FP32 copy and its column panels' running-sum vector.
void require_ready(spk_csr_s& S, bool fp32) {
  const char* msg32 =
      "no-allocation solve: a store's FP32 copy is not made yet (run one FP32 "
      "solve or SpMV with it first)";
  const char* msgp =
      "no-allocation solve: a column-panel store's running-sum vector is not "
      "made yet (run one SpMV with it first)";
  if (S.xt_groups) {
    // window tiles: the tiles' own FP32 values and the long rows' store
    SPK_REQUIRE(!fp32 || (S.xt_has32 && (!S.xt_long || S.xt_long->has32)),
                SPK_INVALID_ARGUMENT, msg32);
    return;
  }
  SPK_REQUIRE(!fp32 || S.has32, SPK_INVALID_ARGUMENT, msg32);
  SPK_REQUIRE(!fp32 || !S.half32 || !S.panels.empty() || S.has16,
              SPK_INVALID_ARGUMENT,
              "no-allocation solve: a store's FP16 copy is not made yet");
  if (!S.panels.empty()) {
    SPK_REQUIRE(fp32 ? S.pacc32.n >= (size_t)S.nrows : S.pacc64.n >= (size_t)S.nrows,
                SPK_INVALID_ARGUMENT, msgp);
    for (spk_csr_s* q : S.panels) SPK_REQUIRE(!fp32 || q->has32, SPK_INVALID_ARGUMENT, msg32);
  }
}

size_t max_blocks(const spk_csr_s& S) {
  if (S.xt_groups) return (size_t)xtile_partials_count(S);
  size_t m = (size_t)S.nblocks;
  for (const spk_csr_s* q : S.panels) m = std::max(m, (size_t)q->nblocks);
  return m;
}

void require_noalloc_precond(spk_precond_s* M, bool fp32, size_t* blocks) {
  if (is_identity(M)) return;
  SPK_REQUIRE(M->kind == spk_precond_s::kJacobi || M->kind == spk_precond_s::kSpmv,
              SPK_INVALID_ARGUMENT,
              "no-allocation solve: preconditioners are identity, Jacobi or an "
              "explicit approximate inverse (ILU(0) keeps lazily made buffers)");
  if (M->kind == spk_precond_s::kJacobi) {
    SPK_REQUIRE(!fp32 || M->has32, SPK_INVALID_ARGUMENT,
                "no-allocation solve: the Jacobi preconditioner's FP32 copy is "
                "not made yet (run one FP32 solve with it first)");
  } else {
    require_ready(*M->M, fp32);
    *blocks = std::max(*blocks, max_blocks(*M->M));
  }
}

}  // namespace
}  // namespace spk

extern "C" spk_status spk_cg_solve_noalloc(spk_csr A, spk_precond M_left,
                                           spk_precond M_right,
                                           const spk_solver_config* cfg,
                                           const void* d_b, void* d_x,
                                           spk_summary* summary,
                                           double* history) {
  return guarded([&] {
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(summary, SPK_INVALID_ARGUMENT, "null summary");
    SPK_REQUIRE(cfg->restart > 0, SPK_INVALID_ARGUMENT,
                "restart_iterations must be > 0 (the reference divides by it)");
    *summary = spk_summary{};
    const bool fp32 = cfg->precision == SPK_F32;
    require_ready(*A, fp32);
    size_t blocks = max_blocks(*A);
    require_noalloc_precond(M_left, fp32, &blocks);
    require_noalloc_precond(M_right, fp32, &blocks);
    // what the solve carves: five n-vectors, the residual history, the
    // state, three hand-off partial buffers and the reduction partials
    const int64_t n = A->nrows;
    const size_t vb = arena_round((size_t)(n > 0 ? n : 1) * (fp32 ? 4 : 8));
    const size_t partials =
        std::max(blocks, (size_t)2 * (size_t)vec_grid(n > 0 ? n : 1));
    size_t need = 5 * vb + arena_round(sizeof(double) * partials) +
                  3 * arena_round(sizeof(double) * 2 * (size_t)kProMaxPartials) +
                  arena_round(1024);
    if (cfg->collect_residuals)
      need += arena_round(sizeof(double) *
                          (size_t)std::max<int64_t>(cfg->max_iterations, 1));
    SPK_REQUIRE(need <= scratch_bytes(), SPK_INVALID_ARGUMENT,
                "no-allocation solve: needs " + std::to_string(need) +
                    " bytes of scratch, more than its " +
                    std::to_string(scratch_bytes()) +
                    " (about 1.5M rows in FP64, 3M in FP32)");
    ArenaScope arena(scratch_ptr(), scratch_bytes(), partials);
    if (fp32)
      cg_solve<float>(A, M_left, M_right, cfg, (const float*)d_b, (float*)d_x,
                      summary, history);
    else
      cg_solve<double>(A, M_left, M_right, cfg, (const double*)d_b,
                       (double*)d_x, summary, history);
  });
}
