// C++ drop-in API tests, written like the reference's Catch2 tests
// (sparkit_testing/*_test.cpp) with a minimal CHECK harness (Catch2 is not
// available).  The matrix type below has the reference's accessor names
// (Compressed_row_matrix.hpp:53-76), so Device_csr::from() is exercised the
// way a sparkit user would call it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "sparkit/hip/sparkit_hip.hpp"

namespace {

int g_failures = 0;
int g_checks = 0;

#define CHECK(cond)                                                      \
  do {                                                                   \
    ++g_checks;                                                          \
    if (!(cond)) {                                                       \
      ++g_failures;                                                      \
      std::printf("  CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #cond); \
    }                                                                    \
  } while (0)

#define CHECK_NEAR(a, b, tol) CHECK(std::fabs((a) - (b)) <= (tol))

template <typename E, typename F>
bool throws(F f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

struct Shape {
  std::ptrdiff_t r, c;
  std::ptrdiff_t row() const { return r; }
  std::ptrdiff_t column() const { return c; }
};

// Host CSR with the reference's accessors; built from (row, col, value)
// entries sorted by (row, col).
struct Host_csr {
  Shape s;
  std::vector<std::ptrdiff_t> rp, ci;
  std::vector<double> v;
  Shape shape() const { return s; }
  const std::vector<std::ptrdiff_t>& row_ptr() const { return rp; }
  const std::vector<std::ptrdiff_t>& col_ind() const { return ci; }
  const std::vector<double>& values() const { return v; }
};

struct Entry {
  std::ptrdiff_t r, c;
  double v;
};

Host_csr make_matrix(std::ptrdiff_t nr, std::ptrdiff_t nc, std::vector<Entry> e) {
  std::sort(e.begin(), e.end(), [](const Entry& a, const Entry& b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  Host_csr A{{nr, nc}, std::vector<std::ptrdiff_t>(nr + 1, 0), {}, {}};
  for (const auto& x : e) {
    A.ci.push_back(x.c);
    A.v.push_back(x.v);
    A.rp[x.r + 1]++;
  }
  for (std::ptrdiff_t r = 0; r < nr; ++r) A.rp[r + 1] += A.rp[r];
  return A;
}

Host_csr tridiagonal(std::ptrdiff_t n, double sub, double diag, double sup) {
  std::vector<Entry> e;
  for (std::ptrdiff_t i = 0; i < n; ++i) {
    if (i > 0) e.push_back({i, i - 1, sub});
    e.push_back({i, i, diag});
    if (i + 1 < n) e.push_back({i, i + 1, sup});
  }
  return make_matrix(n, n, e);
}

// poisson_2d(grid, shift) (matgen.hpp:222-265)
Host_csr poisson_2d(std::ptrdiff_t g, double shift) {
  std::vector<Entry> e;
  for (std::ptrdiff_t r = 0; r < g; ++r)
    for (std::ptrdiff_t c = 0; c < g; ++c) {
      std::ptrdiff_t node = r * g + c, deg = 0;
      if (c > 0) { e.push_back({node, node - 1, -1.0}); ++deg; }
      if (c + 1 < g) { e.push_back({node, node + 1, -1.0}); ++deg; }
      if (r > 0) { e.push_back({node, node - g, -1.0}); ++deg; }
      if (r + 1 < g) { e.push_back({node, node + g, -1.0}); ++deg; }
      e.push_back({node, node, static_cast<double>(deg) + shift});
    }
  return make_matrix(g * g, g * g, e);
}

Host_csr nonsymmetric_sample() {  // matgen.hpp:147-164
  return make_matrix(4, 4, {{0, 0, 4}, {0, 1, -1}, {0, 2, 0.5}, {1, 0, -0.5},
                            {1, 1, 4}, {1, 3, -1}, {2, 0, 0.3}, {2, 2, 4},
                            {2, 3, -0.8}, {3, 1, -0.6}, {3, 2, 0.2}, {3, 3, 4}});
}

std::vector<double> host_multiply(const Host_csr& A, const std::vector<double>& x) {
  std::vector<double> y(A.s.r, 0.0);
  for (std::ptrdiff_t i = 0; i < A.s.r; ++i) {
    double s = 0.0;
    for (auto j = A.rp[i]; j < A.rp[i + 1]; ++j) s += A.v[j] * x[A.ci[j]];
    y[i] = s;
  }
  return y;
}

namespace sh = sparkit::hip;

void test_spmv_known_3x3() {
  // sparse_blas_test.cpp:36-54, through the callable concept
  auto A = sh::Device_csr::from(
      make_matrix(3, 3, {{0, 0, 2}, {0, 2, 3}, {1, 1, 4}, {2, 0, 5}, {2, 2, 6}}));
  std::vector<double> x{1, 2, 3}, y(3);
  A(x.begin(), x.end(), y.begin());
  CHECK(y[0] == 11.0 && y[1] == 8.0 && y[2] == 23.0);
  sh::Device_vector<double> dx(x.begin(), x.end());
  auto dy = sh::multiply(A, dx).to_host();
  CHECK(dy == y);
}

void test_cg_cases() {
  // conjugate_gradient_test.cpp:53-79 identity
  auto I = sh::Device_csr::from(
      make_matrix(4, 4, {{0, 0, 1}, {1, 1, 1}, {2, 2, 1}, {3, 3, 1}}));
  std::vector<double> b{1, 2, 3, 4}, x(4, 0.0);
  sh::CGConfig<double> cfg{1e-12, 50, 100, false};
  auto s = sh::conjugate_gradient(b.begin(), b.end(), x.begin(), x.end(), cfg, I,
                                  sh::Device_identity{}, sh::Device_identity{});
  CHECK(s.converged);
  CHECK(s.computed_iterations <= 1);
  for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], b[i], 1e-10);
  // :174-230 grid 16-node, left Jacobi, right Jacobi
  auto Ah = poisson_2d(4, 5.0);
  auto A = sh::Device_csr::from(Ah);
  std::vector<double> xt(16);
  std::iota(xt.begin(), xt.end(), 1.0);
  auto bb = host_multiply(Ah, xt);
  sh::CGConfig<double> c2{1e-10, 50, 200, true};
  auto J = sh::jacobi(A);
  for (int side = 0; side < 2; ++side) {
    std::vector<double> xx(16, 0.0);
    auto s2 = side == 0
                  ? sh::conjugate_gradient(bb.begin(), bb.end(), xx.begin(),
                                           xx.end(), c2, A, J, sh::Device_identity{})
                  : sh::conjugate_gradient(bb.begin(), bb.end(), xx.begin(),
                                           xx.end(), c2, A, sh::Device_identity{}, J);
    CHECK(s2.converged);
    CHECK(static_cast<std::ptrdiff_t>(s2.iteration_residuals.size()) ==
          s2.computed_iterations);
    for (int i = 0; i < 16; ++i) CHECK_NEAR(xx[i], xt[i], 1e-6);
  }
  // :377-400 max iterations
  std::vector<double> x1(16, 0.0);
  sh::CGConfig<double> c3{1e-14, 50, 1, false};
  auto s3 = sh::conjugate_gradient(bb.begin(), bb.end(), x1.begin(), x1.end(), c3,
                                   A, sh::Device_identity{}, sh::Device_identity{});
  CHECK(!s3.converged);
  CHECK(s3.computed_iterations == 1);
  // zero rhs
  std::vector<double> z(16, 0.0), xz(16, 3.0);
  auto s4 = sh::conjugate_gradient(z.begin(), z.end(), xz.begin(), xz.end(), c2,
                                   A, J, sh::Device_identity{});
  CHECK(s4.converged);
  for (double v : xz) CHECK(v == 0.0);
}

void test_gmres_fgmres_bicgstab() {
  auto Ah = nonsymmetric_sample();
  auto A = sh::Device_csr::from(Ah);
  std::vector<double> xt{1, 2, 3, 4};
  auto b = host_multiply(Ah, xt);
  auto F = sh::ilu0(A);
  // ilu_test.cpp:198-232
  {
    std::vector<double> x(4, 0.0);
    sh::Gmres_config<double> cfg{1e-12, 10, 100, false};
    auto s = sh::gmres(b.begin(), b.end(), x.begin(), x.end(), cfg, A, F,
                       sh::Device_identity{});
    CHECK(s.converged);
    for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], xt[i], 1e-8);
  }
  {
    std::vector<double> x(4, 0.0);
    sh::Fgmres_config<double> cfg{1e-12, 10, 100, true};
    auto s = sh::fgmres(b.begin(), b.end(), x.begin(), x.end(), cfg, A, F);
    CHECK(s.converged);
    for (std::size_t k = 1; k < s.iteration_residuals.size(); ++k)
      CHECK(s.iteration_residuals[k] <= s.iteration_residuals[k - 1] + 1e-14);
    for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], xt[i], 1e-8);
  }
  {
    std::vector<double> x(4, 0.0);
    sh::Bicgstab_config<double> cfg{1e-12, 100, false};
    auto s = sh::bicgstab(b.begin(), b.end(), x.begin(), x.end(), cfg, A,
                          sh::Device_identity{}, F);
    CHECK(s.converged);
    for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], xt[i], 1e-8);
  }
  // bicgstab_test.cpp:392-414: max iterations
  {
    auto Gh = poisson_2d(4, 5.0);
    auto G = sh::Device_csr::from(Gh);
    std::vector<double> xt16(16);
    std::iota(xt16.begin(), xt16.end(), 1.0);
    auto b16 = host_multiply(Gh, xt16);
    std::vector<double> x(16, 0.0);
    sh::Bicgstab_config<double> cfg{1e-14, 2, false};
    auto s = sh::bicgstab(b16.begin(), b16.end(), x.begin(), x.end(), cfg, G,
                          sh::Device_identity{}, sh::Device_identity{});
    CHECK(!s.converged);
    CHECK(s.computed_iterations <= 2);
  }
}

void test_device_vectors_and_fp32() {
  auto Ah = tridiagonal(64, -1.0, 4.0, -1.0);
  auto A = sh::Device_csr::from(Ah);
  std::vector<double> xt(64);
  std::iota(xt.begin(), xt.end(), 1.0);
  auto bh = host_multiply(Ah, xt);
  sh::Device_vector<double> b(bh.begin(), bh.end());
  std::vector<double> zero(64, 0.0);
  sh::Device_vector<double> x(zero.begin(), zero.end());
  sh::CGConfig<double> cfg{1e-12, 1000, 1000, false};
  auto s = sh::conjugate_gradient(b, x, cfg, A, sh::jacobi(A), sh::Device_identity{});
  CHECK(s.converged);
  auto xh = x.to_host();
  for (int i = 0; i < 64; ++i) CHECK_NEAR(xh[i], xt[i], 1e-8);
  // T = float: FP32 vectors, FP64 dot products
  std::vector<float> bf(bh.begin(), bh.end()), xf(64, 0.0f);
  sh::CGConfig<float> cf{1e-5f, 1000, 1000, false};
  auto sf = sh::conjugate_gradient(bf.begin(), bf.end(), xf.begin(), xf.end(), cf,
                                   A, sh::jacobi(A), sh::Device_identity{});
  CHECK(sf.converged);
  for (int i = 0; i < 64; ++i) CHECK_NEAR(xf[i], xt[i], 1e-3 * xt[i]);
}

void test_errors() {
  // jacobi_test.cpp:273-282
  auto Z = sh::Device_csr::from(
      make_matrix(3, 3, {{0, 0, 2}, {0, 1, 1}, {1, 0, 1}, {2, 2, 3}}));
  CHECK(throws<std::invalid_argument>([&] { sh::jacobi(Z); }));
  // ilu_test.cpp:274-282
  auto R = sh::Device_csr::from(make_matrix(3, 4, {{0, 0, 1}, {1, 1, 2}, {2, 2, 3}}));
  CHECK(throws<std::invalid_argument>([&] { sh::ilu0(R); }));
  // Shape.cpp:8-10
  CHECK(throws<std::logic_error>([&] {
    sh::Device_csr::from(make_matrix(1, 3, {{0, 0, 1.0}}));
  }));
  // restart_iterations = 0 divides by zero in the reference
  auto A = sh::Device_csr::from(tridiagonal(4, -1, 4, -1));
  std::vector<double> b{1, 2, 3, 4}, x(4, 0.0);
  sh::CGConfig<double> bad{1e-8, 0, 10, false};
  CHECK(throws<std::invalid_argument>([&] {
    sh::conjugate_gradient(b.begin(), b.end(), x.begin(), x.end(), bad, A,
                           sh::Device_identity{}, sh::Device_identity{});
  }));
}

void test_cgs_bicg() {
  // cgs_test.cpp:107-127 / bicg_test.cpp:139-173
  auto Ah = tridiagonal(4, -1.0, 4.0, -1.0);
  auto A = sh::Device_csr::from(Ah);
  std::vector<double> xt{1, 2, 3, 4};
  auto b = host_multiply(Ah, xt);
  std::vector<double> x(4, 0.0);
  sh::Cgs_config<double> cfg{1e-12, 100, false};
  auto s = sh::cgs(b.begin(), b.end(), x.begin(), x.end(), cfg, A,
                   sh::Device_identity{}, sh::Device_identity{});
  CHECK(s.converged);
  for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], xt[i], 1e-8);
  std::fill(x.begin(), x.end(), 0.0);
  auto At = sh::transpose(A);
  auto s2 = sh::bicg(b.begin(), b.end(), x.begin(), x.end(), cfg, A, At,
                     sh::jacobi(A), sh::Device_identity{});
  CHECK(s2.converged);
  for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], xt[i], 1e-8);
}

void test_minres_qmr() {
  // minres_test.cpp:124-144 / qmr_test.cpp:139-173
  auto Ah = tridiagonal(4, -1.0, 4.0, -1.0);
  auto A = sh::Device_csr::from(Ah);
  std::vector<double> xt{1, 2, 3, 4};
  auto b = host_multiply(Ah, xt);
  std::vector<double> x(4, 0.0);
  sh::Minres_config<double> cfg{1e-12, 100, false};
  auto s = sh::minres(b.begin(), b.end(), x.begin(), x.end(), cfg, A,
                      sh::Device_identity{}, sh::Device_identity{});
  CHECK(s.converged);
  for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], xt[i], 1e-8);
  std::fill(x.begin(), x.end(), 0.0);
  auto At = sh::transpose(A);
  sh::Qmr_config<double> qcfg{1e-12, 100, false};
  auto s2 = sh::qmr(b.begin(), b.end(), x.begin(), x.end(), qcfg, A, At,
                    sh::jacobi(A), sh::Device_identity{});
  CHECK(s2.converged);
  CHECK(s2.computed_iterations % 2 == 0);
  for (int i = 0; i < 4; ++i) CHECK_NEAR(x[i], xt[i], 1e-8);
}

void test_rcm_and_matrix_market() {
  // reordering_test.cpp:144-163 (arrow matrix), matrix_market_test.cpp:186-211
  std::vector<int64_t> rp{0, 4, 6, 8, 10}, ci{0, 1, 2, 3, 0, 1, 0, 2, 0, 3};
  std::vector<double> v{4, 1, 1, 1, 1, 4, 1, 4, 1, 4};
  spk_csr h = nullptr;
  sparkit::hip::check(spk_csr_create(4, 4, 10, rp.data(), ci.data(), v.data(), &h),
                      "create");
  sh::Device_csr A(h);
  auto perm = sh::reverse_cuthill_mckee(A);
  std::vector<int> seen(4, 0);
  for (auto p : perm) {
    CHECK(p >= 0 && p < 4);
    if (p >= 0 && p < 4) seen[(size_t)p] += 1;
  }
  for (int k : seen) CHECK(k == 1);
  const std::string path = "/tmp/spk_cpp_api_test.mtx";
  sh::write_matrix_market(path, A);
  auto B = sh::read_matrix_market(path);
  CHECK(B.rows() == 4);
  CHECK(throws<std::runtime_error>([&] { sh::read_matrix_market("/nonexistent/x.mtx"); }));
}

void test_learned_spai_and_batched() {
  // a 9-point anisotropic diffusion matrix from the device generator
  spk_csr g = nullptr;
  sparkit::hip::check(spk_gen_aniso_q1_2d(24, 5, 1.5, 0, &g), "gen");
  sh::Device_csr A(g);
  sh::Learned_spai net(3);
  const double l0 = net.loss(A);
  std::vector<sh::Device_csr> mats;
  for (int s = 0; s < 4; ++s) {
    spk_csr h = nullptr;
    sparkit::hip::check(spk_gen_aniso_q1_2d(24, 100 + s, 1.5, 0, &h), "gen");
    mats.emplace_back(h);
  }
  auto hist = net.train(mats, 6, 2e-3);
  CHECK(hist.size() == 24);
  CHECK(net.loss(A) < l0);
  auto M = net.predict(A);
  std::vector<double> bh(A.rows(), 1.0), xh(A.rows(), 0.0);
  sh::Gmres_config<double> cfg{1e-9, 30, 2000, false};
  auto s = sh::fgmres(bh.begin(), bh.end(), xh.begin(), xh.end(), cfg, A, M);
  CHECK(s.converged);
  CHECK(throws<std::invalid_argument>([&] {
    net.set_parameters(std::vector<float>(3, 0.f));
  }));
  // checkpoint: a loaded model continues training exactly like the saved one
  const std::string ck = "/tmp/spk_test_api_" + std::to_string(::getpid()) + ".spai";
  net.save(ck);
  sh::Learned_spai twin(77);
  twin.load(ck);
  CHECK(twin.parameters() == net.parameters());
  net.train(mats, 1, 2e-3, 1, 9);
  twin.train(mats, 1, 2e-3, 1, 9);
  CHECK(twin.parameters() == net.parameters());
  std::remove(ck.c_str());
  CHECK(throws<std::runtime_error>([&] { twin.load(ck); }));
  // batched PCG: two systems
  auto B = sh::Batched_pcg::diffusion(5, 9, 4, 12);
  CHECK(B.count() == 5);
  std::vector<double> ones(B.rows(), 1.0), zeros(B.rows(), 0.0);
  sh::Device_vector<double> bb(ones.begin(), ones.end()), xb(zeros.begin(), zeros.end());
  spk_batch_config bc{1e-10, 1e-4, 20, 1000, SPK_F32, 0};
  auto bs = B.solve(bc, bb, xb);
  CHECK(bs.converged == 5);
  CHECK(bs.max_relres < 1e-10);
}

}  // namespace

int main() {
  sparkit::hip::check(spk_init(0), "spk_init");
  struct Case {
    const char* name;
    void (*fn)();
  } cases[] = {{"spmv_known_3x3", test_spmv_known_3x3},
               {"cg_cases", test_cg_cases},
               {"gmres_fgmres_bicgstab", test_gmres_fgmres_bicgstab},
               {"device_vectors_and_fp32", test_device_vectors_and_fp32},
               {"errors", test_errors},
               {"cgs_bicg", test_cgs_bicg},
               {"minres_qmr", test_minres_qmr},
               {"rcm_and_matrix_market", test_rcm_and_matrix_market},
               {"learned_spai_and_batched", test_learned_spai_and_batched}};
  for (const auto& c : cases) {
    int before = g_failures;
    try {
      c.fn();
    } catch (const std::exception& e) {
      ++g_failures;
      std::printf("  unexpected exception: %s\n", e.what());
    }
    std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", c.name);
    std::fflush(stdout);
  }
  std::printf("%d checks, %d failures\n", g_checks, g_failures);
  spk_finalize();
  return g_failures == 0 ? 0 : 1;
}
