// spk_gmres.hip -- device GMRES(m) (gmres.hpp:99-375) and FGMRES(m)
// (fgmres.hpp:96-428).
//
// One restart cycle is queued by the host without reading anything back:
//   r0 = b - A x (SpMV epilogue); GMRES: V0 = M_L r0; beta = |V0|
//   Arnoldi step j:  Z_j = M_R V_j (GMRES) / M V_j (FGMRES)
//                    V_{j+1} = M_L A Z_j (GMRES) / A Z_j (FGMRES)
//                    modified Gram-Schmidt: j+1 fused kernels, kernel i
//                    applies the update of V_{i-1} and reduces <V_i, V_{j+1}>
//                    (so each MGS sweep streams V_{j+1} once per basis
//                    vector, in the reference's order)
//                    last kernel: update with V_j and |V_{j+1}|^2; its last
//                    block runs the Givens rotations and the convergence test
//                    on the device (FP64), then V_{j+1} is normalised
//   form_solution:   back-substitution (one thread) and x += Z y
// Optional orthogonalization SPK_ORTHO_CGS2 (classical Gram-Schmidt applied
// twice): per step two passes of {all j+1 dot products in one launch (a
// chunk x basis-vector grid), one fixed-order finisher, one fused update
// w -= sum_i h_i V_i}, the second update fused with |w|^2 -- nine launches
// per step whatever j is, instead of the 2(j+2)+1 of the sequential MGS
// sweep; each update runs in FP64 and rounds once.
// The host reads the state once per cycle.  Vectors may be FP32
// (precision = SPK_F32); H, the rotations, g, y, every dot product and the
// Gram-Schmidt updates are computed in FP64 and the basis is stored in T.
#include "spk_solver_common.hpp"

namespace spk {
namespace {

struct GmState {
  double bnorm, tol, resnorm, hnorm;
  int64_t iters;
  int32_t stop;       // this cycle is finished (converged / breakdown)
  int32_t done;       // the solve is finished
  int32_t converged;
  int32_t collect;
  int32_t j_final;
  int32_t flexible;
  double* history;
  double* H;  // (m+1) x m, column-major, leading dimension m+1
  double* cs;
  double* sn;
  double* g;
  double* y;
  int64_t m;
};

// apply_givens_rotations + residual record + convergence check
// (gmres.hpp:283-326, 193-205; fgmres.hpp:286-356), thread 0 of the last
// block of the final Gram-Schmidt kernel of step j.
__device__ void arnoldi_finish(GmState* st, int j, double h) {
  const int64_t ld = st->m + 1;
  double* col = st->H + (int64_t)j * ld;
  double* cs = st->cs;
  double* sn = st->sn;
  double* g = st->g;
  col[j + 1] = h;
  st->hnorm = h;
  // the earlier rotations in the reference's order; only the rotated
  // entry carries from one to the next, so each batch of eight loads its
  // rotations and column entries together (one round trip, not eight)
  double ci = col[0];
  for (int i0 = 0; i0 < j; i0 += 8) {
    const int nb = j - i0 < 8 ? j - i0 : 8;
    double c8[8], s8[8], h8[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u < nb) {
        c8[u] = cs[i0 + u];
        s8[u] = sn[i0 + u];
        h8[u] = col[i0 + u + 1];
      }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u < nb) {
        const double temp = c8[u] * ci + s8[u] * h8[u];
        ci = -s8[u] * ci + c8[u] * h8[u];
        col[i0 + u] = temp;
      }
  }
  if (j > 0) col[j] = ci;
  double a = col[j], b = col[j + 1];
  if (b == 0.0) {
    cs[j] = 1.0;
    sn[j] = 0.0;
  } else if (a == 0.0) {
    cs[j] = 0.0;
    sn[j] = 1.0;
  } else {
    double r = sqrt(a * a + b * b);
    cs[j] = a / r;
    sn[j] = b / r;
  }
  col[j] = cs[j] * col[j] + sn[j] * col[j + 1];
  col[j + 1] = 0.0;
  double temp = cs[j] * g[j] + sn[j] * g[j + 1];
  g[j + 1] = -sn[j] * g[j] + cs[j] * g[j + 1];
  g[j] = temp;
  st->iters += 1;
  st->j_final = j;
  double absres = fabs(g[j + 1]);
  st->resnorm = absres;
  if (st->collect) st->history[st->iters - 1] = absres;
  if (absres / st->bnorm < st->tol) {
    st->converged = 1;
    st->stop = 1;
  }
}

constexpr int kDotChunk = 4096;  // elements per (chunk, vector) block

// partial <V_i, w> over one chunk: grid (chunks, j + 1)
template <typename T>
__global__ __launch_bounds__(256) void multi_dot_kernel(
    int64_t n, const T* __restrict__ V, size_t ldv, const T* __restrict__ w,
    int nchunks, double* __restrict__ partials, const int* stop) {
  __shared__ double red[4];
  if (*stop) return;
  const int i = blockIdx.y, c = blockIdx.x;
  const T* __restrict__ vi = V + (size_t)i * ldv;
  const int64_t e0 = (int64_t)c * kDotChunk;
  const int64_t e1 = e0 + kDotChunk < n ? e0 + kDotChunk : n;
  double acc = 0.0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256)
    acc += (double)vi[e] * (double)w[e];
  acc = block_reduce_sum<double, 256>(acc, red);
  if (threadIdx.x == 0) partials[(size_t)i * nchunks + c] = acc;
}

// coef[i] = sum of the chunk partials (fixed order); also added to add_to[i]
__global__ __launch_bounds__(256) void multi_finish_kernel(
    int nchunks, const double* __restrict__ partials, double* __restrict__ coef,
    double* __restrict__ add_to, const int* stop) {
  __shared__ double red[4];
  if (*stop) return;
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int k = threadIdx.x; k < nchunks; k += 256)
    acc += partials[(size_t)i * nchunks + k];
  acc = block_reduce_sum<double, 256>(acc, red);
  if (threadIdx.x == 0) {
    coef[i] = acc;
    if (add_to) add_to[i] += acc;
  }
}

// w -= sum_{i <= j} coef[i] V_i in FP64, one rounding; NORM: partial |w|^2
template <typename T, bool NORM>
__global__ __launch_bounds__(256) void multi_update_kernel(
    int64_t n, const T* __restrict__ V, size_t ldv, int jp1,
    const double* __restrict__ coef, T* __restrict__ w,
    double* __restrict__ partials, const int* stop) {
  extern __shared__ double hc[];
  __shared__ double red[4];
  if (*stop) return;
  for (int i = threadIdx.x; i < jp1; i += 256) hc[i] = coef[i];
  __syncthreads();
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * 256) {
    double s = (double)w[e];
    for (int i = 0; i < jp1; ++i) s -= hc[i] * (double)V[(size_t)i * ldv + e];
    const T t = (T)s;
    w[e] = t;
    if (NORM) acc += (double)t * (double)t;
  }
  if (NORM) {
    acc = block_reduce_sum<double, 256>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
  }
}

// form_solution's back-substitution R y = g (gmres.hpp:333-342), one
// thread in the reference's order.  Up to kBackLds columns the block first
// stages the triangle and g in LDS, so the serial solve reads no global
// memory (a 50-column solve from global memory costs ~125 us of dependent
// load latency).
constexpr int kBackLds = 64;
__global__ __launch_bounds__(256) void back_substitute_kernel(GmState* st) {
  __shared__ double R[kBackLds * kBackLds];
  __shared__ double gs[kBackLds], ys[kBackLds];
  if (st->done) return;
  const int64_t ld = st->m + 1;
  const int nj = st->j_final + 1;
  const double* H = st->H;
  double* y = st->y;
  if (nj > kBackLds) {
    if (threadIdx.x != 0) return;
    for (int ii = nj; ii > 0; --ii) {
      int i = ii - 1;
      double yi = st->g[i];
      for (int k = i + 1; k < nj; ++k) yi -= H[i + (int64_t)k * ld] * y[k];
      y[i] = yi / H[i + (int64_t)i * ld];
    }
    return;
  }
  for (int t = threadIdx.x; t < nj * nj; t += blockDim.x) {
    const int i = t % nj, k = t / nj;  // column k, row i: consecutive in H
    R[i * kBackLds + k] = H[i + (int64_t)k * ld];
  }
  for (int t = threadIdx.x; t < nj; t += blockDim.x) gs[t] = st->g[t];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int ii = nj; ii > 0; --ii) {
      int i = ii - 1;
      double yi = gs[i];
      for (int k = i + 1; k < nj; ++k) yi -= R[i * kBackLds + k] * ys[k];
      ys[i] = yi / R[i * kBackLds + i];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nj; t += blockDim.x) y[t] = ys[t];
}

template <typename T>
void gmres_solve(bool flexible, spk_csr_s* A, spk_precond_s* ML,
                 spk_precond_s* MR, const spk_solver_config* cfg, const T* b,
                 T* x, spk_summary* out, double* history,
                 const CycleMonitor* monitor = nullptr) {
  SPK_REQUIRE(cfg->restart > 0, SPK_INVALID_ARGUMENT,
              "restart_dimension must be > 0");
  const int64_t n = A->nrows, m = cfg->restart, ld = m + 1;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  Workspace& ws = workspace();
  T* V = (T*)ws.get(kWsGmres + 0, (size_t)(m + 1) * nn * sizeof(T));
  // Z is needed whenever the right slot is not the identity
  const bool right_id = flexible ? is_identity(ML) : is_identity(MR);
  T* Z = right_id ? V : (T*)ws.get(kWsGmres + 1, (size_t)m * nn * sizeof(T));
  T* r0 = (T*)ws.get(kWsGmres + 2, nn * sizeof(T));
  double* small = (double*)ws.get(
      kWsGmres + 3, sizeof(double) * (size_t)((m + 1) * m + 3 * m + 2 * (m + 1)));
  double* dhist = cfg->collect_residuals
                      ? (double*)ws.get(kWsGmres + 4, sizeof(double) *
                                        (size_t)std::max<int64_t>(cfg->max_iterations, 1))
                      : nullptr;
  SPK_REQUIRE(cfg->orthogonalization == SPK_ORTHO_MGS ||
                  cfg->orthogonalization == SPK_ORTHO_CGS2 ||
                  cfg->orthogonalization == SPK_ORTHO_CGS1,
              SPK_INVALID_ARGUMENT, "unknown orthogonalization");
  // classical Gram-Schmidt, twice (CGS2) or once (CGS1)
  const bool cgs2 = cfg->orthogonalization != SPK_ORTHO_MGS;
  const int npass = cfg->orthogonalization == SPK_ORTHO_CGS1 ? 1 : 2;
  SPK_REQUIRE(!cgs2 || m < 4096, SPK_INVALID_ARGUMENT,
              "classical Gram-Schmidt supports restart dimensions below 4096");
  double* ctmp = (double*)ws.get(kWsGmres + 6, sizeof(double) * (size_t)(m + 1));
  double* parts = nullptr;
  if (cgs2) {  // sized once: grows (and synchronises) only here
    const size_t nch = (size_t)((n + kDotChunk - 1) / kDotChunk);
    parts = reduce_partials(std::max<size_t>((size_t)(m + 1) * nch,
                                             (size_t)vec_grid(n)));
  }
  DevState<GmState> st(kWsGmres + 5);
  GmState h{};
  h.tol = cfg->tolerance;
  h.collect = cfg->collect_residuals ? 1 : 0;
  h.flexible = flexible;
  h.history = dhist;
  h.H = small;
  h.cs = small + (m + 1) * m;
  h.sn = h.cs + m;
  h.g = h.sn + m;
  h.y = h.g + (m + 1);
  h.m = m;
  st.upload(h);
  GmState* S = st.d;
  hipStream_t s = stream();
  const int* stop = &S->stop;
  const int* done = &S->done;
  spk_precond_s* left = flexible ? nullptr : ML;
  spk_precond_s* right = flexible ? ML : MR;

  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)b[i] * (double)b[i];
  }, [=] __device__(const double* t) { S->bnorm = sqrt(t[0]); });
  h = st.download();
  if (h.bnorm == 0.0) {  // gmres.hpp:165-170
    launch_map(n, [=] __device__(int64_t i) { x[i] = T(0); });
    *out = spk_summary{0.0, 0, 1, SPK_OK, 0};
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return;
  }

  int64_t iters = 0;
  while (iters < cfg->max_iterations && !h.converged && !h.done) {
    TraceRange cycle_("gmres restart cycle");
    // compute_initial_residual (gmres.hpp:211-241 / fgmres.hpp:206-224)
    if (is_identity(left)) {
      launch_spmv(*A, (const T*)x, EpiResidual<T>{b, V}, s);
    } else {
      launch_spmv(*A, (const T*)x, EpiResidual<T>{b, r0}, s);
      left->apply((const T*)r0, V);
    }
    launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
      acc[0] += (double)V[i] * (double)V[i];
    }, [=] __device__(const double* t) {
      double beta = sqrt(t[0]);
      for (int64_t k = 0; k <= S->m; ++k) S->g[k] = 0.0;
      for (int64_t k = 0; k < (S->m + 1) * S->m; ++k) S->H[k] = 0.0;
      for (int64_t k = 0; k < S->m; ++k) S->cs[k] = S->sn[k] = 0.0;
      S->g[0] = beta;
      S->hnorm = beta;
      S->j_final = 0;
      S->stop = 0;
      if (beta == 0.0) {  // x is exact; the reference would divide by zero
        S->resnorm = 0.0;
        S->converged = 1;
        S->stop = 1;
        S->done = 1;
      }
    });
    if (flexible) {
      launch_map(n, [=] __device__(int64_t i) {
        V[i] = (T)((double)V[i] * (1.0 / S->hnorm));
      }, stop);
    } else {
      launch_map(n, [=] __device__(int64_t i) {
        V[i] = (T)((double)V[i] / S->hnorm);
      }, stop);
    }
    const int64_t m_eff = std::min(m, cfg->max_iterations - iters);
    // a preconditioner that runs a solve of its own (inner PCG, inner
    // mixed-precision solve) costs far more than a state read: stop the
    // cycle as soon as it has converged
    const bool check_each = !right_id && right->runs_solve();
    for (int64_t j = 0; j < m_eff; ++j) {
      if (check_each && j > 0 && st.download().stop) break;
      T* vj = V + (size_t)j * nn;
      T* vj1 = V + (size_t)(j + 1) * nn;
      T* zj = right_id ? vj : Z + (size_t)j * nn;
      if (!right_id) right->apply((const T*)vj, zj);
      if (is_identity(left)) {
        launch_spmv(*A, (const T*)zj, EpiStore<T>{vj1, stop}, s);
      } else {
        launch_spmv(*A, (const T*)zj, EpiStore<T>{r0, stop}, s);
        left->apply((const T*)r0, vj1);
      }
      if (cgs2) {
        double* hcol = small + j * ld;
        const int jp1 = (int)(j + 1);
        const int nch = (int)((n + kDotChunk - 1) / kDotChunk);
        const size_t dyn = sizeof(double) * (size_t)jp1;
        for (int pass = 0; pass < npass; ++pass) {
          hipLaunchKernelGGL((multi_dot_kernel<T>), dim3(nch, jp1), dim3(256), 0, s,
                             n, (const T*)V, nn, (const T*)vj1, nch, parts, stop);
          hipLaunchKernelGGL(multi_finish_kernel, dim3(jp1), dim3(256), 0, s, nch,
                             (const double*)parts, pass == 0 ? hcol : ctmp,
                             pass == 0 ? (double*)nullptr : hcol, stop);
          if (pass + 1 < npass) {
            hipLaunchKernelGGL((multi_update_kernel<T, false>), dim3(vec_grid(n)),
                               dim3(256), dyn, s, n, (const T*)V, nn, jp1,
                               (const double*)hcol, vj1, parts, stop);
          } else {
            const int g = vec_grid(n);
            hipLaunchKernelGGL((multi_update_kernel<T, true>), dim3(g), dim3(256),
                               dyn, s, n, (const T*)V, nn, jp1,
                               (const double*)(pass == 0 ? hcol : ctmp), vj1, parts,
                               stop);
            const int jj = (int)j;
            launch_finish<1>(parts, g, [=] __device__(const double* t) {
              arnoldi_finish(S, jj, sqrt(t[0]));
            }, stop);
          }
          SPK_HIP_TRY(hipGetLastError());
        }
      } else {
      // modified Gram-Schmidt (gmres.hpp:258-267)
      for (int64_t i = 0; i <= j; ++i) {
        const T* vi = V + (size_t)i * nn;
        const T* vim1 = i > 0 ? V + (size_t)(i - 1) * nn : nullptr;
        double* hcol = small + j * ld;
        const int64_t ii = i;
        launch_reduce<1>(n, [=] __device__(int64_t e, double* acc) {
          T w = vj1[e];
          if (ii > 0) {
            w = (T)((double)w - hcol[ii - 1] * (double)vim1[e]);
            vj1[e] = w;
          }
          acc[0] += (double)vi[e] * (double)w;
        }, [=] __device__(const double* t) { hcol[ii] = t[0]; }, stop);
      }
      {
        const double* hcol = small + j * ld;
        const int jj = (int)j;
        launch_reduce<1>(n, [=] __device__(int64_t e, double* acc) {
          T w = (T)((double)vj1[e] - hcol[jj] * (double)vj[e]);
          vj1[e] = w;
          acc[0] += (double)w * (double)w;
        }, [=] __device__(const double* t) {
          arnoldi_finish(S, jj, sqrt(t[0]));
        }, stop);
      }
      }  // MGS
      // normalise V_{j+1} if h > 0 (gmres.hpp:276-280 divides,
      // fgmres.hpp:283 multiplies by the reciprocal)
      if (flexible) {
        launch_map(n, [=] __device__(int64_t e) {
          double hn = S->hnorm;
          if (hn > 0.0) vj1[e] = (T)((double)vj1[e] * (1.0 / hn));
        }, stop);
      } else {
        launch_map(n, [=] __device__(int64_t e) {
          double hn = S->hnorm;
          if (hn > 0.0) vj1[e] = (T)((double)vj1[e] / hn);
        }, stop);
      }
    }
    // form_solution (gmres.hpp:328-354)
    hipLaunchKernelGGL(back_substitute_kernel, dim3(1), dim3(256), 0, s, S);
    const T* Zc = Z;
    launch_map(n, [=] __device__(int64_t e) {
      const int nj = S->j_final + 1;
      double sum = 0.0;
      for (int k = 0; k < nj; ++k) sum += (double)Zc[(size_t)k * nn + e] * S->y[k];
      x[e] = (T)((double)x[e] + sum);
    }, done);
    // close the cycle: a converged cycle finishes the solve
    launch_map(1, [=] __device__(int64_t) {
      if (S->converged) S->done = 1;
    }, done);
    h = st.download();
    iters = h.iters;
    if (monitor && !h.converged && !h.done && !(*monitor)(iters, h.resnorm))
      break;
  }
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

void fgmres_monitored(spk_csr_s* A, spk_precond_s* M,
                      const spk_solver_config* cfg, const double* b, double* x,
                      spk_summary* out, const CycleMonitor& monitor) {
  *out = spk_summary{};
  gmres_solve<double>(true, A, M, nullptr, cfg, b, x, out, nullptr, &monitor);
}

}  // namespace spk

using namespace spk;

extern "C" {

spk_status spk_gmres_solve(spk_csr A, spk_precond M_left, spk_precond M_right,
                           const spk_solver_config* cfg, const void* d_b,
                           void* d_x, spk_summary* summary, double* history) {
  return guarded([&] {
    TraceRange trace_("spk_gmres_solve");
    check_solver_args(A, M_left, M_right, cfg, d_b, d_x);
    SPK_REQUIRE(summary, SPK_INVALID_ARGUMENT, "null summary");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      gmres_solve<float>(false, A, M_left, M_right, cfg, (const float*)d_b,
                         (float*)d_x, summary, history);
    else
      gmres_solve<double>(false, A, M_left, M_right, cfg, (const double*)d_b,
                          (double*)d_x, summary, history);
    check_preconds(M_left, M_right);
  });
}

spk_status spk_fgmres_solve(spk_csr A, spk_precond M,
                            const spk_solver_config* cfg, const void* d_b,
                            void* d_x, spk_summary* summary, double* history) {
  return guarded([&] {
    TraceRange trace_("spk_fgmres_solve");
    check_solver_args(A, M, nullptr, cfg, d_b, d_x);
    SPK_REQUIRE(summary, SPK_INVALID_ARGUMENT, "null summary");
    *summary = spk_summary{};
    if (cfg->precision == SPK_F32)
      gmres_solve<float>(true, A, M, nullptr, cfg, (const float*)d_b,
                         (float*)d_x, summary, history);
    else
      gmres_solve<double>(true, A, M, nullptr, cfg, (const double*)d_b,
                          (double*)d_x, summary, history);
    check_preconds(M, nullptr);
  });
}

}  // extern "C"
