// sparkit/hip/sparkit_hip.hpp -- C++ drop-in for sparkit's solve path on
// MI355X, header-only over the C ABI in sparkit_hip.h.
//
// The reference's solvers (namespace sparkit::data::detail) take
//   conjugate_gradient(bfirst, blast, xfirst, xlast, CGConfig<T>,
//                      linear_operator, left_preconditioner,
//                      right_preconditioner) -> CGSummary<T>
// (conjugate_gradient.hpp:339-366), and likewise gmres (gmres.hpp:400-427),
// fgmres (fgmres.hpp:452-470) and bicgstab (bicgstab.hpp:383-410).  This
// header provides the same free functions in sparkit::hip with the operator
// and preconditioner slots filled by DEVICE objects:
//
//   Device_csr       the matrix (Compressed_row_matrix<T>, uploaded once)
//   Device_jacobi    jacobi(A) / jacobi_apply          (jacobi.hpp:25-50)
//   Device_ilu0      ilu0(A) / milu0(A) / ilu_apply    (ilu.hpp:35-179)
//   Device_identity  the identity slot (conjugate_gradient_test.cpp:45-47)
//   Device_spai      z = M r with an explicit approximate inverse M
//
// The config argument is any struct with the reference's field names
// (tolerance, restart_iterations / restart_dimension, max_iterations,
// collect_residuals), so the reference's own CGConfig<double> etc. can be
// passed unchanged; T = double runs in FP64, T = float in FP32 (with FP64
// dot products and scalars).  Summaries have the reference's fields.  b and x
// may be host iterators (copied in once, x copied back once) or
// Device_vector<T>s (no copies).  Status codes become the reference's
// exception types: std::invalid_argument, std::logic_error; other failures
// raise std::runtime_error.  A numerical breakdown the reference leaves
// undetected (BiCGSTAB rho = 0, omega = 0) ends the solve with
// converged = false and breakdown = true.
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "sparkit_hip.h"

namespace sparkit::hip {

using size_type = std::ptrdiff_t;

inline void check(spk_status st, const char* where) {
  if (st == SPK_OK) return;
  std::string msg = std::string(where) + ": " + spk_last_error();
  if (st == SPK_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  if (st == SPK_LOGIC_ERROR) throw std::logic_error(msg);
  throw std::runtime_error(msg);
}

template <typename T>
inline constexpr int dtype_code() {
  static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>,
                "value type must be double or float");
  return std::is_same_v<T, double> ? SPK_F64 : SPK_F32;
}

// ------------------------------------------------------------------ vectors

// A vector in HBM.
template <typename T>
class Device_vector {
 public:
  Device_vector() = default;
  explicit Device_vector(size_type n) : n_(n) {
    void* p = nullptr;
    check(spk_malloc(&p, bytes() ? bytes() : 1), "spk_malloc");
    p_.reset(static_cast<T*>(p), [](T* q) { spk_free(q); });
  }
  template <typename It>
  Device_vector(It first, It last)
      : Device_vector(static_cast<size_type>(std::distance(first, last))) {
    assign(first, last);
  }
  template <typename It>
  void assign(It first, It last) {
    std::vector<T> h(first, last);
    if (static_cast<size_type>(h.size()) != n_)
      throw std::invalid_argument("Device_vector::assign: size mismatch");
    if (n_) check(spk_memcpy_htod(p_.get(), h.data(), bytes()), "spk_memcpy_htod");
  }
  template <typename OutIt>
  void copy_to(OutIt out) const {
    std::vector<T> h(static_cast<std::size_t>(n_));
    if (n_) check(spk_memcpy_dtoh(h.data(), p_.get(), bytes()), "spk_memcpy_dtoh");
    std::copy(h.begin(), h.end(), out);
  }
  std::vector<T> to_host() const {
    std::vector<T> h(static_cast<std::size_t>(n_));
    copy_to(h.begin());
    return h;
  }
  T* data() const { return p_.get(); }
  size_type size() const { return n_; }
  std::size_t bytes() const { return static_cast<std::size_t>(n_) * sizeof(T); }

 private:
  size_type n_ = 0;
  std::shared_ptr<T> p_;
};

// -------------------------------------------------------------- the matrix

// Compressed_row_matrix<double> resident in HBM (int32 indices, FP64 values
// plus an FP32 copy made on first FP32 use).  Copies share the store, so
// passing it by value into a solver (conjugate_gradient.hpp:299-301) is free.
class Device_csr {
 public:
  Device_csr(size_type rows, size_type cols,
             const std::vector<std::ptrdiff_t>& row_ptr,
             const std::vector<std::ptrdiff_t>& col_ind,
             const std::vector<double>& values)
      : Device_csr(rows, cols, row_ptr.data(), col_ind.data(), values.data(),
                   static_cast<size_type>(col_ind.size())) {}
  Device_csr(size_type rows, size_type cols, const std::ptrdiff_t* row_ptr,
             const std::ptrdiff_t* col_ind, const double* values,
             size_type nnz) {
    static_assert(sizeof(std::ptrdiff_t) == sizeof(int64_t));
    spk_csr h = nullptr;
    check(spk_csr_create(rows, cols, nnz,
                         reinterpret_cast<const int64_t*>(row_ptr),
                         reinterpret_cast<const int64_t*>(col_ind), values, &h),
          "spk_csr_create");
    h_.reset(h, [](spk_csr q) { spk_csr_destroy(q); });
    rows_ = rows;
    cols_ = cols;
    nnz_ = nnz;
  }
  explicit Device_csr(spk_csr adopted) {
    h_.reset(adopted, [](spk_csr q) { spk_csr_destroy(q); });
    int64_t r = 0, c = 0, z = 0;
    check(spk_csr_shape(adopted, &r, &c, &z), "spk_csr_shape");
    rows_ = r;
    cols_ = c;
    nnz_ = z;
  }
  // From any matrix with the reference's accessors: shape().row(),
  // shape().column(), row_ptr(), col_ind(), values()
  // (Compressed_row_matrix.hpp:53-76).
  template <typename Matrix>
  static Device_csr from(const Matrix& A) {
    auto rp = A.row_ptr();
    auto ci = A.col_ind();
    auto v = A.values();
    return Device_csr(A.shape().row(), A.shape().column(), rp.data(),
                      ci.data(), v.data(), static_cast<size_type>(ci.size()));
  }

  spk_csr handle() const { return h_.get(); }
  size_type rows() const { return rows_; }
  size_type cols() const { return cols_; }
  size_type size() const { return nnz_; }

  // multiply(A, x) on device vectors (sparse_blas.hpp:17-34)
  template <typename T>
  void multiply(const Device_vector<T>& x, Device_vector<T>& y) const {
    check(spk_csr_spmv(h_.get(), dtype_code<T>(), x.data(), y.data()),
          "spk_csr_spmv");
  }
  // The reference's callable concept op(first, last, out) on host iterators
  // (one upload and one download per call).
  template <typename InIt, typename OutIt>
  void operator()(InIt first, InIt last, OutIt out) const {
    using T = std::remove_cv_t<typename std::iterator_traits<InIt>::value_type>;
    Device_vector<T> x(first, last), y(rows_);
    multiply(x, y);
    y.copy_to(out);
  }

 private:
  std::shared_ptr<spk_csr_s> h_;
  size_type rows_ = 0, cols_ = 0, nnz_ = 0;
};

// multiply(A, x) returning a new vector, as sparse_blas.hpp:17-34 does
template <typename T>
Device_vector<T> multiply(const Device_csr& A, const Device_vector<T>& x) {
  Device_vector<T> y(A.rows());
  A.multiply(x, y);
  return y;
}

// ------------------------------------------------------- preconditioners

class Device_preconditioner {
 public:
  spk_precond handle() const { return h_.get(); }
  size_type size() const { return n_; }
  template <typename T>
  void apply(const Device_vector<T>& r, Device_vector<T>& z) const {
    check(spk_precond_apply(h_.get(), dtype_code<T>(), r.data(), z.data()),
          "spk_precond_apply");
  }
  template <typename InIt, typename OutIt>
  void operator()(InIt first, InIt last, OutIt out) const {
    using T = std::remove_cv_t<typename std::iterator_traits<InIt>::value_type>;
    Device_vector<T> r(first, last), z(r.size());
    apply(r, z);
    z.copy_to(out);
  }

 protected:
  void adopt(spk_precond p, size_type n) {
    h_.reset(p, [](spk_precond q) { spk_precond_destroy(q); });
    n_ = n;
  }
  std::shared_ptr<spk_precond_s> h_;
  size_type n_ = 0;
};

// jacobi(A): throws std::invalid_argument on a zero diagonal (jacobi.hpp:31)
class Device_jacobi : public Device_preconditioner {
 public:
  explicit Device_jacobi(const Device_csr& A) {
    spk_precond p = nullptr;
    check(spk_precond_create_jacobi(A.handle(), &p), "jacobi");
    adopt(p, A.rows() < A.cols() ? A.rows() : A.cols());
  }
};

// ilu0(A) / milu0(A) (ilu.hpp:155-168): std::invalid_argument for a
// non-square matrix or a missing diagonal
class Device_ilu0 : public Device_preconditioner {
 public:
  explicit Device_ilu0(const Device_csr& A, bool modified = false) {
    spk_precond p = nullptr;
    check(spk_precond_create_ilu0(A.handle(), modified ? 1 : 0, &p), "ilu0");
    adopt(p, A.rows());
  }
  // 0: exact triangular solves (ilu_apply); k > 0: k Jacobi sweeps per
  // triangle (an approximate, bandwidth-bound apply)
  void set_jacobi_sweeps(int k) {
    check(spk_ilu0_set_jacobi_sweeps(h_.get(), k), "spk_ilu0_set_jacobi_sweeps");
  }
  int jacobi_sweeps() const {
    int k = 0;
    check(spk_ilu0_get_jacobi_sweeps(h_.get(), &k), "spk_ilu0_get_jacobi_sweeps");
    return k;
  }
};

// z = M r with an explicit sparse approximate inverse
class Device_spai : public Device_preconditioner {
 public:
  explicit Device_spai(const Device_csr& M) : M_(M) {
    spk_precond p = nullptr;
    check(spk_precond_create_spmv(M.handle(), &p), "spai");
    adopt(p, M.rows());
  }

 private:
  Device_csr M_;  // keeps the store alive
};

// The learned sparse approximate inverse: a network trained on the device
// (bf16 MFMA kernels, Adam) to minimise ||I - M A||_F^2 that predicts M on
// A's pattern; predict() gives a Device_spai for the preconditioner slot of
// fgmres (fgmres.hpp:452-470).  A new network predicts the Jacobi inverse.
class Learned_spai {
 public:
  static constexpr int parameter_count = SPK_SPAI_NPARAM;
  explicit Learned_spai(uint64_t seed = 0) {
    spk_spai_model m = nullptr;
    check(spk_spai_model_create(seed, &m), "spk_spai_model_create");
    h_.reset(m, [](spk_spai_model q) { spk_spai_model_destroy(q); });
  }
  // Adam over the matrices; returns each step's mean loss
  std::vector<double> train(const std::vector<Device_csr>& mats, int64_t epochs,
                            double learning_rate = 1e-3, int64_t batch = 1,
                            uint64_t shuffle_seed = 0) {
    std::vector<spk_csr> hs;
    for (const auto& m : mats) hs.push_back(m.handle());
    spk_spai_train_config c{epochs, learning_rate, batch, shuffle_seed};
    const int64_t steps =
        epochs * ((static_cast<int64_t>(mats.size()) + batch - 1) / batch);
    std::vector<double> hist(static_cast<size_t>(steps > 0 ? steps : 1));
    check(spk_spai_train(h_.get(), hs.data(), static_cast<int64_t>(hs.size()), &c,
                         hist.data()),
          "spk_spai_train");
    hist.resize(static_cast<size_t>(steps));
    return hist;
  }
  double loss(const Device_csr& A, std::vector<float>* grad = nullptr) const {
    double l = 0.0;
    if (grad) grad->resize(parameter_count);
    check(spk_spai_loss_grad(h_.get(), A.handle(), &l, grad ? grad->data() : nullptr),
          "spk_spai_loss_grad");
    return l;
  }
  std::vector<float> parameters() const {
    std::vector<float> p(parameter_count);
    check(spk_spai_model_get_params(h_.get(), p.data()), "spk_spai_model_get_params");
    return p;
  }
  void set_parameters(const std::vector<float>& p) {
    if (p.size() != static_cast<size_t>(parameter_count))
      throw std::invalid_argument("Learned_spai: wrong parameter count");
    check(spk_spai_model_set_params(h_.get(), p.data()), "spk_spai_model_set_params");
  }
  // training loss: SPK_SPAI_LOSS_INVERSE (||I - M A||, the default) or
  // SPK_SPAI_LOSS_SYMMETRIC (||I - S A||, S = (M + M^T) / 2)
  void set_objective(int objective) {
    check(spk_spai_model_set_objective(h_.get(), objective),
          "spk_spai_model_set_objective");
  }
  // the training state (weights, Adam moments and step) to a checkpoint
  // file and back; file errors throw std::runtime_error
  void save(const std::string& path) const {
    check(spk_spai_model_save(h_.get(), path.c_str()), "spk_spai_model_save");
  }
  void load(const std::string& path) {
    check(spk_spai_model_load(h_.get(), path.c_str()), "spk_spai_model_load");
  }
  Device_csr approximate_inverse(const Device_csr& A) const {
    spk_csr M = nullptr;
    check(spk_spai_predict(h_.get(), A.handle(), &M), "spk_spai_predict");
    return Device_csr(M);
  }
  Device_spai predict(const Device_csr& A) const {
    return Device_spai(approximate_inverse(A));
  }

 private:
  std::shared_ptr<spk_spai_model_s> h_;
};

// Many independent SPD systems (the diagonal blocks of one store) solved by
// batched Jacobi-PCG in one persistent launch (FP32 inner solve with an FP64
// true-residual refinement loop, or FP64 throughout).
class Batched_pcg {
 public:
  static Batched_pcg diffusion(int64_t count, uint64_t seed, int64_t side_min,
                               int64_t side_max) {
    spk_batch b = nullptr;
    check(spk_batch_create_diffusion(count, seed, side_min, side_max, &b),
          "spk_batch_create_diffusion");
    return Batched_pcg(b);
  }
  Batched_pcg(const Device_csr& A, const std::vector<int64_t>& row_offsets) {
    spk_batch b = nullptr;
    check(spk_batch_create(A.handle(), static_cast<int64_t>(row_offsets.size()) - 1,
                           row_offsets.data(), &b),
          "spk_batch_create");
    adopt(b);
  }
  int64_t count() const { return count_; }
  int64_t rows() const { return rows_; }
  spk_batch_tuning autotune(const spk_batch_config& c, const Device_vector<double>& b,
                            Device_vector<double>& x, int64_t sample = 512) {
    spk_batch_tuning t{};
    check(spk_batch_autotune(h_.get(), &c, b.data(), x.data(), sample, &t, nullptr,
                             nullptr, nullptr),
          "spk_batch_autotune");
    return t;
  }
  void set_tuning(const spk_batch_tuning& t) {
    check(spk_batch_set_tuning(h_.get(), &t), "spk_batch_set_tuning");
  }
  spk_batch_summary solve(const spk_batch_config& c, const Device_vector<double>& b,
                          Device_vector<double>& x, bool use_graph = true) {
    spk_batch_summary s{};
    check(spk_batch_pcg_solve(h_.get(), &c, b.data(), x.data(), use_graph ? 1 : 0, &s),
          "spk_batch_pcg_solve");
    return s;
  }

 private:
  explicit Batched_pcg(spk_batch b) { adopt(b); }
  void adopt(spk_batch b) {
    h_.reset(b, [](spk_batch q) { spk_batch_destroy(q); });
    check(spk_batch_info(b, &count_, &rows_, nullptr, nullptr), "spk_batch_info");
  }
  std::shared_ptr<spk_batch_s> h_;
  int64_t count_ = 0, rows_ = 0;
};

struct Device_identity {
  spk_precond handle() const { return nullptr; }
  template <typename InIt, typename OutIt>
  void operator()(InIt first, InIt last, OutIt out) const {
    std::copy(first, last, out);
  }
};

inline Device_jacobi jacobi(const Device_csr& A) { return Device_jacobi(A); }
inline Device_ilu0 ilu0(const Device_csr& A) { return Device_ilu0(A, false); }
inline Device_ilu0 milu0(const Device_csr& A) { return Device_ilu0(A, true); }

// --------------------------------------------------------- configs/summaries

// Field-for-field mirrors (conjugate_gradient.hpp:26-74, gmres.hpp:24-68,
// fgmres.hpp:24-68, bicgstab.hpp:24-63); the solvers accept these or the
// reference's own structs.
template <typename T>
struct CGConfig {
  T tolerance{};
  size_type restart_iterations{};
  size_type max_iterations{};
  bool collect_residuals{};
};
template <typename T>
struct Gmres_config {
  T tolerance{};
  size_type restart_dimension{};
  size_type max_iterations{};
  bool collect_residuals{};
  // not in the reference: SPK_ORTHO_MGS (its algorithm), SPK_ORTHO_CGS2 or
  // SPK_ORTHO_CGS1
  int orthogonalization{SPK_ORTHO_MGS};
};
template <typename T>
using Fgmres_config = Gmres_config<T>;
template <typename T>
struct Bicgstab_config {
  T tolerance{};
  size_type max_iterations{};
  bool collect_residuals{};
};

template <typename T>
struct Summary {
  T residual_norm{};
  size_type computed_iterations{};
  bool converged{};
  std::vector<T> iteration_residuals{};
  bool breakdown{};  // not in the reference: a detected breakdown
};
template <typename T>
using CGSummary = Summary<T>;
template <typename T>
using Gmres_summary = Summary<T>;
template <typename T>
using Fgmres_summary = Summary<T>;
template <typename T>
using Bicgstab_summary = Summary<T>;

namespace detail {

template <typename Cfg>
using value_of = std::remove_cv_t<decltype(std::declval<Cfg>().tolerance)>;

template <typename P>
spk_precond slot(const P& p) {
  return p.handle();
}

template <typename T>
Summary<T> to_summary(const spk_summary& s, const std::vector<double>& h) {
  Summary<T> out;
  out.residual_norm = static_cast<T>(s.residual_norm);
  out.computed_iterations = s.computed_iterations;
  out.converged = s.converged != 0;
  out.breakdown = s.status == SPK_BREAKDOWN;
  for (int64_t k = 0; k < s.history_len; ++k)
    out.iteration_residuals.push_back(static_cast<T>(h[static_cast<std::size_t>(k)]));
  return out;
}

template <typename T, typename Solve>
Summary<T> run_device(const spk_solver_config& c, Solve solve) {
  spk_summary s{};
  std::vector<double> hist(static_cast<std::size_t>(c.max_iterations + 1));
  check(solve(&c, &s, hist.data()), "solve");
  return to_summary<T>(s, hist);
}

// Host iterators: copy b and x in, solve on the device, copy x back.
template <typename T, typename BIter, typename XIter, typename DevSolve>
Summary<T> run_host(BIter bfirst, BIter blast, XIter xfirst, XIter xlast,
                    DevSolve dev) {
  Device_vector<T> b(bfirst, blast), x(xfirst, xlast);
  Summary<T> s = dev(b, x);
  x.copy_to(xfirst);
  return s;
}

template <typename T>
spk_solver_config make_config(double tol, size_type restart,
                              size_type max_it, bool collect) {
  spk_solver_config c{};
  c.tolerance = tol;
  c.restart = restart;
  c.max_iterations = max_it;
  c.collect_residuals = collect ? 1 : 0;
  c.precision = dtype_code<T>();
  return c;
}

// orthogonalization of a config that has the field (this layer's
// Gmres_config); the reference's own structs get MGS, its algorithm
template <typename Config>
int ortho_of(const Config& c) {
  if constexpr (requires { c.orthogonalization; })
    return static_cast<int>(c.orthogonalization);
  else
    return SPK_ORTHO_MGS;
}

}  // namespace detail

// ------------------------------------------------------------------ solvers

// conjugate_gradient (conjugate_gradient.hpp:339-366), device vectors
template <typename Config, typename Left, typename Right,
          typename T = detail::value_of<Config>>
Summary<T> conjugate_gradient(const Device_vector<T>& b, Device_vector<T>& x,
                              Config cfg, const Device_csr& A,
                              const Left& M_left, const Right& M_right) {
  auto c = detail::make_config<T>(cfg.tolerance, cfg.restart_iterations,
                                  cfg.max_iterations, cfg.collect_residuals);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_cg_solve(A.handle(), detail::slot(M_left),
                        detail::slot(M_right), pc, b.data(), x.data(), s, h);
  });
}

// conjugate_gradient with the reference's iterator signature
template <typename BIter, typename XIter, typename Config, typename Left,
          typename Right, typename T = detail::value_of<Config>>
Summary<T> conjugate_gradient(BIter bfirst, BIter blast, XIter xfirst,
                              XIter xlast, Config cfg, const Device_csr& A,
                              const Left& M_left, const Right& M_right) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return conjugate_gradient(b, x, cfg, A, M_left,
                                                         M_right);
                             });
}

// gmres (gmres.hpp:400-427)
template <typename Config, typename Left, typename Right,
          typename T = detail::value_of<Config>>
Summary<T> gmres(const Device_vector<T>& b, Device_vector<T>& x, Config cfg,
                 const Device_csr& A, const Left& M_left,
                 const Right& M_right) {
  auto c = detail::make_config<T>(cfg.tolerance, cfg.restart_dimension,
                                  cfg.max_iterations, cfg.collect_residuals);
  c.orthogonalization = detail::ortho_of(cfg);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_gmres_solve(A.handle(), detail::slot(M_left),
                           detail::slot(M_right), pc, b.data(), x.data(), s, h);
  });
}

template <typename BIter, typename XIter, typename Config, typename Left,
          typename Right, typename T = detail::value_of<Config>>
Summary<T> gmres(BIter bfirst, BIter blast, XIter xfirst, XIter xlast,
                 Config cfg, const Device_csr& A, const Left& M_left,
                 const Right& M_right) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return gmres(b, x, cfg, A, M_left, M_right);
                             });
}

// fgmres (fgmres.hpp:452-470)
template <typename Config, typename Prec,
          typename T = detail::value_of<Config>>
Summary<T> fgmres(const Device_vector<T>& b, Device_vector<T>& x, Config cfg,
                  const Device_csr& A, const Prec& M) {
  auto c = detail::make_config<T>(cfg.tolerance, cfg.restart_dimension,
                                  cfg.max_iterations, cfg.collect_residuals);
  c.orthogonalization = detail::ortho_of(cfg);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_fgmres_solve(A.handle(), detail::slot(M), pc, b.data(),
                            x.data(), s, h);
  });
}

template <typename BIter, typename XIter, typename Config, typename Prec,
          typename T = detail::value_of<Config>>
Summary<T> fgmres(BIter bfirst, BIter blast, XIter xfirst, XIter xlast,
                  Config cfg, const Device_csr& A, const Prec& M) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return fgmres(b, x, cfg, A, M);
                             });
}

// bicgstab (bicgstab.hpp:383-410); computed_iterations counts matvecs
template <typename Config, typename Left, typename Right,
          typename T = detail::value_of<Config>>
Summary<T> bicgstab(const Device_vector<T>& b, Device_vector<T>& x, Config cfg,
                    const Device_csr& A, const Left& M_left,
                    const Right& M_right) {
  auto c = detail::make_config<T>(cfg.tolerance, 0, cfg.max_iterations,
                                  cfg.collect_residuals);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_bicgstab_solve(A.handle(), detail::slot(M_left),
                              detail::slot(M_right), pc, b.data(), x.data(),
                              s, h);
  });
}

template <typename BIter, typename XIter, typename Config, typename Left,
          typename Right, typename T = detail::value_of<Config>>
Summary<T> bicgstab(BIter bfirst, BIter blast, XIter xfirst, XIter xlast,
                    Config cfg, const Device_csr& A, const Left& M_left,
                    const Right& M_right) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return bicgstab(b, x, cfg, A, M_left, M_right);
                             });
}

// cgs (cgs.hpp:383-395); Cgs_config has Bicgstab_config's fields
template <typename T>
using Cgs_config = Bicgstab_config<T>;
template <typename Config, typename Left, typename Right,
          typename T = detail::value_of<Config>>
Summary<T> cgs(const Device_vector<T>& b, Device_vector<T>& x, Config cfg,
               const Device_csr& A, const Left& M_left, const Right& M_right) {
  auto c = detail::make_config<T>(cfg.tolerance, 0, cfg.max_iterations,
                                  cfg.collect_residuals);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_cgs_solve(A.handle(), detail::slot(M_left), detail::slot(M_right),
                         pc, b.data(), x.data(), s, h);
  });
}

template <typename BIter, typename XIter, typename Config, typename Left,
          typename Right, typename T = detail::value_of<Config>>
Summary<T> cgs(BIter bfirst, BIter blast, XIter xfirst, XIter xlast, Config cfg,
               const Device_csr& A, const Left& M_left, const Right& M_right) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return cgs(b, x, cfg, A, M_left, M_right);
                             });
}

// A^T as a store: the transpose operator of bicg (multiply_transpose,
// sparse_blas.hpp:36-52)
inline Device_csr transpose(const Device_csr& A) {
  spk_csr t = nullptr;
  check(spk_csr_transpose(A.handle(), &t), "spk_csr_transpose");
  return Device_csr(t);
}

// read_matrix_market + to_compressed_row / write_matrix_market
// (matrix_market.hpp:43-160); failures throw std::runtime_error
inline Device_csr read_matrix_market(const std::string& path) {
  spk_csr t = nullptr;
  check(spk_csr_read_matrix_market(path.c_str(), &t), "read_matrix_market");
  return Device_csr(t);
}
inline void write_matrix_market(const std::string& path, const Device_csr& A) {
  check(spk_csr_write_matrix_market(A.handle(), path.c_str()), "write_matrix_market");
}

// reverse_cuthill_mckee(A.sparsity()) (reordering.cpp:152-203): perm[old] = new
inline std::vector<int64_t> reverse_cuthill_mckee(const Device_csr& A) {
  Device_vector<double> d(A.rows());  // n 8-byte slots for the device result
  check(spk_csr_rcm(A.handle(), reinterpret_cast<int64_t*>(d.data())),
        "reverse_cuthill_mckee");
  std::vector<int64_t> perm(A.rows());
  check(spk_memcpy_dtoh(perm.data(), d.data(), perm.size() * sizeof(int64_t)),
        "reverse_cuthill_mckee");
  return perm;
}

// bicg (bicg.hpp:86-397): the transpose operator slot takes the store of
// A^T; Bicg_config has Bicgstab_config's fields
template <typename T>
using Bicg_config = Bicgstab_config<T>;
template <typename Config, typename Left, typename Right,
          typename T = detail::value_of<Config>>
Summary<T> bicg(const Device_vector<T>& b, Device_vector<T>& x, Config cfg,
                const Device_csr& A, const Device_csr& At, const Left& M_left,
                const Right& M_right) {
  auto c = detail::make_config<T>(cfg.tolerance, 0, cfg.max_iterations,
                                  cfg.collect_residuals);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_bicg_solve(A.handle(), At.handle(), detail::slot(M_left),
                          detail::slot(M_right), pc, b.data(), x.data(), s, h);
  });
}

template <typename BIter, typename XIter, typename Config, typename Left,
          typename Right, typename T = detail::value_of<Config>>
Summary<T> bicg(BIter bfirst, BIter blast, XIter xfirst, XIter xlast, Config cfg,
                const Device_csr& A, const Device_csr& At, const Left& M_left,
                const Right& M_right) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return bicg(b, x, cfg, A, At, M_left, M_right);
                             });
}

// minres (minres.hpp:364-395); Minres_config has Bicgstab_config's fields
template <typename T>
using Minres_config = Bicgstab_config<T>;
template <typename Config, typename Left, typename Right,
          typename T = detail::value_of<Config>>
Summary<T> minres(const Device_vector<T>& b, Device_vector<T>& x, Config cfg,
                  const Device_csr& A, const Left& M_left, const Right& M_right) {
  auto c = detail::make_config<T>(cfg.tolerance, 0, cfg.max_iterations,
                                  cfg.collect_residuals);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_minres_solve(A.handle(), detail::slot(M_left),
                            detail::slot(M_right), pc, b.data(), x.data(), s, h);
  });
}

template <typename BIter, typename XIter, typename Config, typename Left,
          typename Right, typename T = detail::value_of<Config>>
Summary<T> minres(BIter bfirst, BIter blast, XIter xfirst, XIter xlast,
                  Config cfg, const Device_csr& A, const Left& M_left,
                  const Right& M_right) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return minres(b, x, cfg, A, M_left, M_right);
                             });
}

// qmr (qmr.hpp:455-489): the transpose operator slot takes the store of
// A^T; Qmr_config has Bicgstab_config's fields
template <typename T>
using Qmr_config = Bicgstab_config<T>;
template <typename Config, typename Left, typename Right,
          typename T = detail::value_of<Config>>
Summary<T> qmr(const Device_vector<T>& b, Device_vector<T>& x, Config cfg,
               const Device_csr& A, const Device_csr& At, const Left& M_left,
               const Right& M_right) {
  auto c = detail::make_config<T>(cfg.tolerance, 0, cfg.max_iterations,
                                  cfg.collect_residuals);
  return detail::run_device<T>(c, [&](const spk_solver_config* pc,
                                      spk_summary* s, double* h) {
    return spk_qmr_solve(A.handle(), At.handle(), detail::slot(M_left),
                         detail::slot(M_right), pc, b.data(), x.data(), s, h);
  });
}

template <typename BIter, typename XIter, typename Config, typename Left,
          typename Right, typename T = detail::value_of<Config>>
Summary<T> qmr(BIter bfirst, BIter blast, XIter xfirst, XIter xlast, Config cfg,
               const Device_csr& A, const Device_csr& At, const Left& M_left,
               const Right& M_right) {
  return detail::run_host<T>(bfirst, blast, xfirst, xlast,
                             [&](Device_vector<T>& b, Device_vector<T>& x) {
                               return qmr(b, x, cfg, A, At, M_left, M_right);
                             });
}

}  // namespace sparkit::hip
