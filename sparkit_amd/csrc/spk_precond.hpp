// spk_precond.hpp -- device preconditioners: identity, Jacobi, ILU(0)/MILU(0)
// and "apply an explicit sparse approximate inverse" (one SpMV).
//
// ILU(0) layout in HBM:
//   L store: strictly-lower entries of A's pattern + unit diagonal LAST in
//            each row; U store: diagonal FIRST + upper entries (the
//            reference's layout, ilu.hpp:123-141), int32 indices.
//   lvl_rows_l / lvl_ptr_l: rows of the forward sweep (and of the numeric
//            factorization) grouped by dependency level;
//   lvl_rows_u / lvl_ptr_u: the same for the backward sweep.
// A level is one launch; rows in a level are independent.  The numeric
// factorization runs the reference's IKJ loop (ilu.hpp:72-115) for every row
// of a level in one thread, with contraction of multiply-adds disabled, so
// the factors are bit-identical to the sequential reference.
#pragma once

#include "spk_csr.hpp"

struct spk_precond_s {
  enum Kind { kIdentity, kJacobi, kIlu, kSpmv, kInnerCg, kInnerIr } kind = kIdentity;
  int64_t n = 0;
  // Jacobi
  spk::DevBuf<double> inv_d64;
  spk::DevBuf<float> inv_d32;
  bool has32 = false;
  // ILU(0)
  spk_csr_s* L = nullptr;
  spk_csr_s* U = nullptr;
  spk::DevBuf<int32_t> lvl_rows_l, lvl_rows_u;
  std::vector<int32_t> lvl_ptr_l, lvl_ptr_u;
  // the same orders with every level padded by -1 entries to a multiple of
  // the wavefront size (the sweeps' thread -> row map)
  spk::DevBuf<int32_t> sweep_rows_l, sweep_rows_u;
  int64_t sweep_len_l = 0, sweep_len_u = 0;
  unsigned sweep_grid_l = 1, sweep_grid_u = 1;  // persistent sweep grids
  spk::DevBuf<double> tmp64;  // between the sweeps; kept filled with the
  spk::DevBuf<float> tmp32;   // ready sentinel between applications
  spk::DevBuf<int> sweep_error;  // set if a sweep's bounded wait expired
  // Sweep schedule, chosen when the factors are made: one launch per
  // dependency level when the levels are wide (average at least
  // kLevelLaunchWidth rows: thousands of waiting wavefronts would flood the
  // memory system with readiness polls), the sync-free sweep otherwise
  // (long chains of narrow levels, where launches would dominate).
  bool level_launch = false;
  // Approximate apply (jac_sweeps > 0): each triangular solve is replaced
  // by jac_sweeps Jacobi iterations from the diagonal-scaled right-hand
  // side, y <- r - L_s y and z <- (y - U_s z) / d_U, with L_s, U_s the
  // strict parts (a fixed linear operator: k terms of the Neumann series;
  // exact once k reaches the dependency depth, the strict parts being
  // nilpotent).  The strict stores and d_U are made on first use.
  int jac_sweeps = 0;
  spk_csr_s* Ls = nullptr;
  spk_csr_s* Us = nullptr;
  spk::DevBuf<double> du64, jb64a, jb64b, jb64c;
  spk::DevBuf<float> du32, jb32a, jb32b, jb32c;
  bool du_has32 = false;
  // SPAI
  spk_csr_s* M = nullptr;  // not owned
  // Inner PCG (a Krylov preconditioner for FGMRES's flexible slot): z = the
  // iterate of up to inner_iters PCG steps on inner_A z = r from z = 0 with
  // inner_M in its preconditioner slot, in inner_precision, stopping at a
  // relative residual of inner_tol.  inner_A and inner_M are not owned.
  spk_csr_s* inner_A = nullptr;
  spk_precond_s* inner_M = nullptr;
  int64_t inner_iters = 0;
  double inner_tol = 0.0;
  int inner_precision = 1;  // SPK_F32
  int64_t inner_total = 0;  // inner steps run by all applications so far
  spk::DevBuf<float> ib32, iz32;
  spk::DevBuf<double> ib64, iz64;
  // Inner mixed-precision solve (kInnerIr, also for FGMRES's flexible
  // slot): z = the result of spk_ir_solve on inner_A z = r from z = 0 with
  // inner_M and the configuration inner_ir.
  spk_ir_config inner_ir{};
  // a preconditioner that runs a solve of its own: not a fixed operator
  bool runs_solve() const { return kind == kInnerCg || kind == kInnerIr; }

  ~spk_precond_s();
  const float* inv_diag32();
  template <typename T>
  const T* inv_diag();
  template <typename T>
  void apply(const T* in, T* out);
  // standalone sweeps (x is filled with the sentinel first)
  template <typename T>
  void forward(const T* b, T* x);
  template <typename T>
  void backward(const T* b, T* x);
  template <typename T>
  T* tmp();
  template <typename T, bool LOWER>
  void level_sweep(const T* b, T* x);
  // throws if a triangular sweep reported an expired wait (synchronises)
  void check_sweeps();
  // refill the scratch vectors with the ready sentinel
  void reset_sentinels();
  // the strict parts and U's diagonal for the approximate apply
  void build_strict();
  template <typename T>
  void jacobi_apply(const T* in, T* out);
  template <typename T>
  void inner_cg_apply(const T* in, T* out);
  template <typename T>
  void inner_ir_apply(const T* in, T* out);
};

template <>
inline const double* spk_precond_s::inv_diag<double>() { return inv_d64.p; }
template <>
inline const float* spk_precond_s::inv_diag<float>() { return inv_diag32(); }
