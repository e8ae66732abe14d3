// spk_blas1.hip -- BLAS-1 entry points, the FP64 true-residual check and the
// device generators (stencil matrices written straight into CSR, seeded
// right-hand sides).
#include <algorithm>
#include <cmath>
#include <vector>

#include "spk_csr.hpp"
#include "spk_vec.hpp"

namespace spk {
namespace {

template <typename T>
double dot_impl(int64_t n, const T* x, const T* y) {
  DevBuf<double> res(1);
  double* r = res.p;
  launch_reduce<1>(n, [=] __device__(int64_t i, double* acc) {
    acc[0] += (double)x[i] * (double)y[i];
  }, [=] __device__(const double* t) { *r = t[0]; });
  double h = 0.0;
  SPK_HIP_TRY(hipMemcpyAsync(&h, r, sizeof(double), hipMemcpyDeviceToHost, stream()));
  SPK_HIP_TRY(hipStreamSynchronize(stream()));
  return h;
}

struct EpiResidualSq {  // accumulates (b_i - (A x)_i)^2
  static constexpr bool kReduce = true;
  const double* __restrict__ b;
  double* out;
  const int* done = nullptr;
  __device__ double row(int i, double s) const {
    double r = b[i] - s;
    return r * r;
  }
  __device__ void finalize(double t) const { *out = t; }
};

// Row lengths of the stencil generators; the fill kernels then write each
// row's columns in ascending order.
__global__ void poisson2d_len_kernel(int64_t nx, int64_t ny, int32_t* len) {
  int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (node >= nx * ny) return;
  int64_t r = node / nx, c = node % nx;
  len[node] = 1 + (c > 0) + (c + 1 < nx) + (r > 0) + (r + 1 < ny);
}

// poisson_2d(nx, ny, shift), matgen.hpp:173-258: off-diagonals -1, diagonal
// degree + shift; make_matrix sorts each row, so the columns come out as
// node-nx, node-1, node, node+1, node+nx.
__global__ void poisson2d_fill_kernel(int64_t nx, int64_t ny, double shift,
                                      const int32_t* rp, int32_t* ci,
                                      double* v) {
#pragma clang fp contract(off)
  int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (node >= nx * ny) return;
  int64_t r = node / nx, c = node % nx;
  int k = rp[node];
  int degree = (c > 0) + (c + 1 < nx) + (r > 0) + (r + 1 < ny);
  if (r > 0) { ci[k] = (int32_t)(node - nx); v[k++] = -1.0; }
  if (c > 0) { ci[k] = (int32_t)(node - 1); v[k++] = -1.0; }
  ci[k] = (int32_t)node;
  v[k++] = (double)degree + shift;
  if (c + 1 < nx) { ci[k] = (int32_t)(node + 1); v[k++] = -1.0; }
  if (r + 1 < ny) { ci[k] = (int32_t)(node + nx); v[k++] = -1.0; }
}

__global__ void stencil3d_len_kernel(int64_t nx, int64_t ny, int64_t nz,
                                     int32_t* len) {
  int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= nx * ny * nz) return;
  int64_t i = row % nx, j = (row / nx) % ny, k = row / (nx * ny);
  len[row] = 1 + (i > 0) + (i + 1 < nx) + (j > 0) + (j + 1 < ny) + (k > 0) +
             (k + 1 < nz);
}

// 7-point 3D operator: off-diagonal values (lower-side and upper-side
// neighbour per axis) and a fixed diagonal.
struct Stencil7 {
  double diag;
  double xm, xp, ym, yp, zm, zp;
};

__global__ void stencil3d_fill_kernel(int64_t nx, int64_t ny, int64_t nz,
                                      Stencil7 st, const int32_t* rp,
                                      int32_t* ci, double* v) {
  int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= nx * ny * nz) return;
  int64_t plane = nx * ny;
  int64_t i = row % nx, j = (row / nx) % ny, k = row / plane;
  int p = rp[row];
  if (k > 0) { ci[p] = (int32_t)(row - plane); v[p++] = st.zm; }
  if (j > 0) { ci[p] = (int32_t)(row - nx); v[p++] = st.ym; }
  if (i > 0) { ci[p] = (int32_t)(row - 1); v[p++] = st.xm; }
  ci[p] = (int32_t)row;
  v[p++] = st.diag;
  if (i + 1 < nx) { ci[p] = (int32_t)(row + 1); v[p++] = st.xp; }
  if (j + 1 < ny) { ci[p] = (int32_t)(row + nx); v[p++] = st.yp; }
  if (k + 1 < nz) { ci[p] = (int32_t)(row + plane); v[p++] = st.zp; }
}

__device__ inline uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Allocate a store of n rows whose row lengths `len_kernel` writes, take the
// prefix sum on the host (it is needed there for the plan anyway) and let
// `fill` write columns and values.
template <typename LenK, typename FillK>
spk_csr_s* build_generated(int64_t n, LenK len_kernel, FillK fill) {
  SPK_REQUIRE(n > 0 && n < INT32_MAX, SPK_INVALID_ARGUMENT, "bad grid size");
  hipStream_t s = stream();
  auto* A = new spk_csr_s();
  try {
    DevBuf<int32_t> len((size_t)n);
    unsigned g = (unsigned)((n + 255) / 256);
    len_kernel(g, len.p);
    std::vector<int32_t> hl((size_t)n);
    SPK_HIP_TRY(hipMemcpyAsync(hl.data(), len.p, n * 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    std::vector<int32_t> rp((size_t)n + 1);
    int64_t acc = 0;
    rp[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
      acc += hl[r];
      SPK_REQUIRE(acc < INT32_MAX, SPK_INVALID_ARGUMENT, "nnz exceeds int32");
      rp[r + 1] = (int32_t)acc;
    }
    A->allocate(n, n, acc);
    SPK_HIP_TRY(hipMemcpyAsync(A->rp.p, rp.data(), (n + 1) * 4,
                               hipMemcpyHostToDevice, s));
    fill(g, A->rp.p, A->ci.p, A->v64.p);
    SPK_HIP_TRY(hipGetLastError());
    finish_store(*A, false);
  } catch (...) {
    delete A;
    throw;
  }
  return A;
}

// 2D anisotropic variable-coefficient diffusion, bilinear (Q1) elements on
// the unit square with the Dirichlet boundary eliminated (N x N interior
// nodes, 9-point stencil).  The per-cell coefficients and the 4x4 element
// table are drawn on the host with libm (exp, log, cos, sin, sqrt):
//   sP = seed ^ 0x414E49: eps = exp(ln(1000) u01(mix(sP, 0))),
//                         theta = pi u01(mix(sP, 1));
//   K0 = R(theta) diag(1, eps) R(theta)^T;
//   E[a][b] = ((kxx X[a][b] + kyy Y[a][b]) + kxy Z[a][b]) / 6 for the corner
//   order q = dx + 2 dy (X, Y, Z: 6 x the exact integrals of the products of
//   the bilinear basis gradients over the unit square);
//   sK = seed ^ 0x4B4150: kappa(cell id) = exp(sigma z), z from Box-Muller
//   with u1 = 1 - u01(mix(sK, 2 id)), u2 = u01(mix(sK, 2 id + 1)).
// The device assembles A(p, p') = sum over the shared cells, ascending cell
// id, of kappa * E[q_p][q_p'] (the same order on every path, so A is
// bitwise symmetric).
struct Q1Table {
  double e[16];
};

inline uint64_t host_mix(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline double host_u01(uint64_t z) {
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

Q1Table aniso_q1_params(int64_t N, uint64_t seed, double sigma, bool iso,
                        std::vector<double>& kappa) {
#pragma clang fp contract(off)
  static const double X[4][4] = {
      {2, -2, 1, -1}, {-2, 2, -1, 1}, {1, -1, 2, -2}, {-1, 1, -2, 2}};
  static const double Y[4][4] = {
      {2, 1, -2, -1}, {1, 2, -1, -2}, {-2, -1, 2, 1}, {-1, -2, 1, 2}};
  static const double Z[4][4] = {
      {3, 0, 0, -3}, {0, -3, 3, 0}, {0, 3, -3, 0}, {-3, 0, 0, 3}};
  double eps = 1.0, theta = 0.0;
  if (!iso) {
    eps = std::exp(6.907755278982137 * host_u01(host_mix(seed ^ 0x414E49ULL, 0)));
    theta = 3.141592653589793 * host_u01(host_mix(seed ^ 0x414E49ULL, 1));
  }
  const double c = std::cos(theta), s = std::sin(theta);
  const double kxx = c * c + eps * (s * s);
  const double kyy = s * s + eps * (c * c);
  const double kxy = (1.0 - eps) * (c * s);
  Q1Table t;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      t.e[4 * a + b] = ((kxx * X[a][b] + kyy * Y[a][b]) + kxy * Z[a][b]) / 6.0;
  const int64_t nc = N + 1;
  kappa.assign((size_t)(nc * nc), 1.0);
  if (sigma != 0.0)
    for (int64_t id = 0; id < nc * nc; ++id) {
      double u1 = 1.0 - host_u01(host_mix(seed ^ 0x4B4150ULL, 2 * (uint64_t)id));
      double u2 = host_u01(host_mix(seed ^ 0x4B4150ULL, 2 * (uint64_t)id + 1));
      double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      kappa[(size_t)id] = std::exp(sigma * z);
    }
  return t;
}

__global__ void q1_len_kernel(int64_t N, int32_t* len) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= N * N) return;
  int64_t ix = p % N + 1, iy = p / N + 1;
  int cx = 3 - (ix == 1) - (ix == N), cy = 3 - (iy == 1) - (iy == N);
  len[p] = cx * cy;
}

__global__ void q1_fill_kernel(int64_t N, Q1Table t,
                               const double* __restrict__ kappa,
                               const int32_t* __restrict__ rp,
                               int32_t* __restrict__ ci, double* __restrict__ v) {
#pragma clang fp contract(off)
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= N * N) return;
  const int64_t ix = p % N + 1, iy = p / N + 1, nc = N + 1;
  int32_t k = rp[p];
  for (int oy = -1; oy <= 1; ++oy)
    for (int ox = -1; ox <= 1; ++ox) {
      const int64_t jx = ix + ox, jy = iy + oy;
      if (jx < 1 || jx > N || jy < 1 || jy > N) continue;
      double s = 0.0;
      for (int b = 0; b <= 1; ++b)
        for (int a = 0; a <= 1; ++a) {
          const int64_t cx = ix - 1 + a, cy = iy - 1 + b;
          const int64_t dxq = jx - cx, dyq = jy - cy;
          if (dxq < 0 || dxq > 1 || dyq < 0 || dyq > 1) continue;
          const int qp = (int)((ix - cx) + 2 * (iy - cy));
          const int qq = (int)(dxq + 2 * dyq);
          s += kappa[cy * nc + cx] * t.e[4 * qp + qq];
        }
      ci[k] = (int32_t)((jy - 1) * N + (jx - 1));
      v[k] = s;
      ++k;
    }
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

spk_status spk_dot(int dtype, int64_t n, const void* d_x, const void* d_y,
                   double* result) {
  return guarded([&] {
    SPK_REQUIRE(result && n >= 0, SPK_INVALID_ARGUMENT, "bad argument");
    SPK_REQUIRE(dtype == SPK_F32 || dtype == SPK_F64, SPK_INVALID_ARGUMENT,
                "dtype must be SPK_F32 or SPK_F64");
    SPK_REQUIRE(n == 0 || (d_x && d_y), SPK_INVALID_ARGUMENT,
                "null vector pointer");
    *result = dtype == SPK_F32 ? dot_impl(n, (const float*)d_x, (const float*)d_y)
                               : dot_impl(n, (const double*)d_x, (const double*)d_y);
  });
}

spk_status spk_axpby(int dtype, int64_t n, double a, const void* d_x, double b,
                     void* d_y) {
  return guarded([&] {
    SPK_REQUIRE(n >= 0, SPK_INVALID_ARGUMENT, "bad argument");
    SPK_REQUIRE(dtype == SPK_F32 || dtype == SPK_F64, SPK_INVALID_ARGUMENT,
                "dtype must be SPK_F32 or SPK_F64");
    SPK_REQUIRE(n == 0 || (d_x && d_y), SPK_INVALID_ARGUMENT,
                "null vector pointer");
    if (dtype == SPK_F32) {
      const float* x = (const float*)d_x;
      float* y = (float*)d_y;
      float fa = (float)a, fb = (float)b;
      launch_map(n, [=] __device__(int64_t i) { y[i] = fa * x[i] + fb * y[i]; });
    } else {
      const double* x = (const double*)d_x;
      double* y = (double*)d_y;
      launch_map(n, [=] __device__(int64_t i) { y[i] = a * x[i] + b * y[i]; });
    }
  });
}

spk_status spk_convert(int64_t n, int dtype_src, const void* d_src,
                       int dtype_dst, void* d_dst) {
  return guarded([&] {
    if (dtype_src == SPK_F64 && dtype_dst == SPK_F32) {
      const double* a = (const double*)d_src;
      float* b = (float*)d_dst;
      launch_map(n, [=] __device__(int64_t i) { b[i] = (float)a[i]; });
    } else if (dtype_src == SPK_F32 && dtype_dst == SPK_F64) {
      const float* a = (const float*)d_src;
      double* b = (double*)d_dst;
      launch_map(n, [=] __device__(int64_t i) { b[i] = (double)a[i]; });
    } else {
      size_t bytes = (size_t)n * (dtype_src == SPK_F32 ? 4 : 8);
      SPK_HIP_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice,
                                 stream()));
    }
  });
}

spk_status spk_true_relres(spk_csr A, const double* d_b, const double* d_x,
                           double* relres) {
  return guarded([&] {
    SPK_REQUIRE(A && d_b && d_x && relres, SPK_INVALID_ARGUMENT, "null argument");
    DevBuf<double> out(2);
    double* o = out.p;
    launch_spmv(*A, d_x, EpiResidualSq{d_b, o}, stream());
    launch_reduce<1>(A->nrows, [=] __device__(int64_t i, double* acc) {
      acc[0] += d_b[i] * d_b[i];
    }, [=] __device__(const double* t) { o[1] = t[0]; });
    double h[2];
    SPK_HIP_TRY(hipMemcpyAsync(h, o, sizeof(h), hipMemcpyDeviceToHost, stream()));
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    *relres = std::sqrt(h[0]) / std::sqrt(h[1]);
  });
}

spk_status spk_fill_uniform_pm1(uint64_t seed, int64_t n, double* d_out) {
  return guarded([&] {
    SPK_REQUIRE(d_out && n >= 0, SPK_INVALID_ARGUMENT, "bad argument");
    launch_map(n, [=] __device__(int64_t k) {
#pragma clang fp contract(off)
      double u = (double)(splitmix64_at(seed, (uint64_t)k) >> 11) *
                 (1.0 / 9007199254740992.0);
      d_out[k] = 2.0 * u - 1.0;
    });
  });
}

spk_status spk_gen_poisson_2d(int64_t nx, int64_t ny, double shift,
                              spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out && nx > 0 && ny > 0, SPK_INVALID_ARGUMENT, "bad argument");
    SPK_REQUIRE(nx * ny > 1, SPK_LOGIC_ERROR, "invalid matrix shape");
    hipStream_t s = stream();
    *out = build_generated(
        nx * ny,
        [&](unsigned g, int32_t* len) {
          hipLaunchKernelGGL(poisson2d_len_kernel, dim3(g), dim3(256), 0, s, nx,
                             ny, len);
        },
        [&](unsigned g, const int32_t* rp, int32_t* ci, double* v) {
          hipLaunchKernelGGL(poisson2d_fill_kernel, dim3(g), dim3(256), 0, s,
                             nx, ny, shift, rp, ci, v);
        });
  });
}

static spk_status gen_stencil3d(int64_t nx, int64_t ny, int64_t nz,
                                Stencil7 st, spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out && nx > 0 && ny > 0 && nz > 0, SPK_INVALID_ARGUMENT,
                "bad argument");
    SPK_REQUIRE(nx * ny * nz > 1, SPK_LOGIC_ERROR, "invalid matrix shape");
    hipStream_t s = stream();
    *out = build_generated(
        nx * ny * nz,
        [&](unsigned g, int32_t* len) {
          hipLaunchKernelGGL(stencil3d_len_kernel, dim3(g), dim3(256), 0, s, nx,
                             ny, nz, len);
        },
        [&](unsigned g, const int32_t* rp, int32_t* ci, double* v) {
          hipLaunchKernelGGL(stencil3d_fill_kernel, dim3(g), dim3(256), 0, s,
                             nx, ny, nz, st, rp, ci, v);
        });
  });
}

spk_status spk_gen_poisson_3d(int64_t nx, int64_t ny, int64_t nz,
                              spk_csr* out) {
  Stencil7 st{6.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0};
  return gen_stencil3d(nx, ny, nz, st, out);
}

spk_status spk_gen_convdiff_3d(int64_t nx, int64_t ny, int64_t nz, double d,
                               double cx, double cy, double cz, spk_csr* out) {
  // upwind neighbour -(d + c_k), downwind -d, diagonal ((6d + cx) + cy) + cz
  // (evaluated on the host, in the same order as the CPU restatement)
  volatile double six_d = 6.0 * d;
  double diag = ((six_d + cx) + cy) + cz;
  Stencil7 st{diag, -d - cx, -d, -d - cy, -d, -d - cz, -d};
  return gen_stencil3d(nx, ny, nz, st, out);
}

spk_status spk_gen_aniso_q1_2d(int64_t n_side, uint64_t seed, double sigma,
                               int isotropic, spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out && n_side >= 2 && n_side < 46000 && sigma >= 0.0,
                SPK_INVALID_ARGUMENT, "bad argument");
    hipStream_t s = stream();
    std::vector<double> kappa;
    Q1Table t = aniso_q1_params(n_side, seed, sigma, isotropic != 0, kappa);
    DevBuf<double> dk(kappa.size());
    SPK_HIP_TRY(hipMemcpyAsync(dk.p, kappa.data(), kappa.size() * 8,
                               hipMemcpyHostToDevice, s));
    const int64_t N = n_side;
    *out = build_generated(
        N * N,
        [&](unsigned g, int32_t* len) {
          hipLaunchKernelGGL(q1_len_kernel, dim3(g), dim3(256), 0, s, N, len);
        },
        [&](unsigned g, const int32_t* rp, int32_t* ci, double* v) {
          hipLaunchKernelGGL(q1_fill_kernel, dim3(g), dim3(256), 0, s, N, t,
                             (const double*)dk.p, rp, ci, v);
        });
    SPK_HIP_TRY(hipStreamSynchronize(s));
  });
}

}  // extern "C"
