// spk_batched.hip -- batched Jacobi-PCG over many independent SPD systems in
// one persistent launch.
//
// Storage: the systems are the diagonal blocks of one CSR store (global row
// and column numbering, int32 indices), plus a system table of row offsets.
// Per-row vectors of the whole batch sit side by side in HBM: b, x, r (FP64)
// and the inner solve's r, d, p, q, 1/diag (FP32 or FP64).
//
// Kernel: every workgroup pulls the next system from a queue (a global
// counter, systems ordered by decreasing size so the long solves start
// first) and solves it alone:
//   r = b - A x (FP64); Jacobi-PCG in the inner precision (FP32 values and
//   vectors, dot products accumulated in FP64) with reliable updates: the
//   correction d accumulates in FP32 and, whenever the recurrence residual
//   has dropped by inner_tolerance since the last update, x += d in FP64
//   and the FP64 true residual replaces the recurrence one, keeping the
//   search direction (no restart, so the iteration count stays that of
//   plain PCG).  The solve ends when the FP64 true residual is below tol.
//   With the FP64 inner precision this is plain FP64 Jacobi-PCG judged on
//   the true residual.
// Per inner iteration three passes over the system's rows, each ending in a
// workgroup reduction (wavefront butterfly + LDS, fixed order):
//   q = A p and <p,q>;  d += a p, r -= a q, <r, D^-1 r>, <r,r>;
//   p = D^-1 r + b p.
// Systems with at most lds_rows rows keep p in LDS, so the SpMV gathers
// come from LDS; larger ones gather p through the caches.
// The launch shape (threads per workgroup, workgroups per CU) and lds_rows
// are chosen by an autotuner that times candidate shapes on a stratified
// sample of the batch.  A solve can be captured into a hipGraph (a reset
// kernel for the queue counter and x, then the persistent kernel) and
// replayed.
#include <algorithm>
#include <cmath>
#include <map>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "spk_csr.hpp"
#include "spk_vec.hpp"

namespace spk {
namespace {

// ------------------------------------------------------------ generator
// System k of a batch (seed, sides [smin, smax]), cell-centred finite
// volumes, 5-point, Dirichlet:
//   s_k = smin + mix(seed ^ 0x534944, k) mod (smax - smin + 1)
//   t_k = mix(seed ^ 0x424154, k); cell c = y s + x:
//   kappa_c = exp(sqrt(-2 log(1 - u01(mix(t_k, 2c)))) cos(2 pi u01(mix(t_k, 2c+1))))
//   interior face t = (2 kappa_a) kappa_b / (kappa_a + kappa_b), boundary
//   face t = 2 kappa_c; row c: -t_S, -t_W, ((t_S + t_W) + t_E) + t_N, -t_E,
//   -t_N over the neighbours inside the grid.
// kappa is drawn on the host (libm), the matrix assembled on the device.

uint64_t hmix(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double hu01(uint64_t z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ double face_coef(double a, double b) {
  return (2.0 * a) * b / (a + b);
}

__global__ void bd_fill_kernel(int64_t total, int K,
                               const int64_t* __restrict__ sys_row,
                               const int64_t* __restrict__ sys_nnz,
                               const int32_t* __restrict__ side,
                               const double* __restrict__ kap,
                               int32_t* __restrict__ rp, int32_t* __restrict__ ci,
                               double* __restrict__ v) {
#pragma clang fp contract(off)
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = K - 1;  // system holding row g
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (sys_row[mid] <= g) lo = mid; else hi = mid - 1;
  }
  const int64_t r0 = sys_row[lo];
  const int64_t s = side[lo];
  const int64_t c = g - r0, y = c / s, x = c % s;
  const int64_t rows_before = y == 0 ? 0 : (4 * s - 2) + (y - 1) * (5 * s - 2);
  const int64_t per = 1 + (y > 0) + (y < s - 1);
  const int64_t in_row = x * per + (x > 0 ? x - 1 : 0) + x;
  int64_t p = sys_nnz[lo] + rows_before + in_row;
  rp[g] = (int32_t)p;
  if (g == total - 1) rp[total] = (int32_t)(sys_nnz[lo] + 5 * s * s - 4 * s);
  const double kc = kap[g];
  const double tS = y > 0 ? face_coef(kc, kap[g - s]) : 2.0 * kc;
  const double tW = x > 0 ? face_coef(kc, kap[g - 1]) : 2.0 * kc;
  const double tE = x < s - 1 ? face_coef(kc, kap[g + 1]) : 2.0 * kc;
  const double tN = y < s - 1 ? face_coef(kc, kap[g + s]) : 2.0 * kc;
  if (y > 0) { ci[p] = (int32_t)(g - s); v[p++] = -tS; }
  if (x > 0) { ci[p] = (int32_t)(g - 1); v[p++] = -tW; }
  ci[p] = (int32_t)g;
  v[p++] = ((tS + tW) + tE) + tN;
  if (x < s - 1) { ci[p] = (int32_t)(g + 1); v[p++] = -tE; }
  if (y < s - 1) { ci[p] = (int32_t)(g + s); v[p++] = -tN; }
}

// ------------------------------------------------------------- solver

struct BatchArgs {
  const int32_t* __restrict__ rp;
  const int32_t* __restrict__ ci;
  const double* __restrict__ v64;
  const void* __restrict__ vin;     // inner-precision values
  const int64_t* __restrict__ sys_row;
  const int32_t* __restrict__ order;
  int K;
  const double* __restrict__ b;
  double* __restrict__ x;
  double* __restrict__ r64;
  void* r;   // inner precision vectors
  void* d;
  void* p;
  void* q;
  const void* invd;
  int* counter;
  int32_t* iters;
  int32_t* outer;
  double* relres;
  int32_t* conv;
  double tol, inner_tol;
  int max_outer, max_inner;
  int lds_rows;
  // sliced-ELL copy of the inner-precision matrix (ell_w slots per row,
  // slot-major within each system: slot k of local row i of the system
  // starting at global row r0 (n rows) at r0 * ell_w + k * n + i; short
  // rows padded with (own row, 0)), used when ell_w > 0.  Its columns are
  // either global (ell_ci) or, when every one lies within 32767 of its row,
  // 16-bit offsets from the row (ell_off; then ell_ci is null)
  const int32_t* __restrict__ ell_ci;
  const int16_t* __restrict__ ell_off;
  const void* __restrict__ ell_v;
  int64_t ell_n;
  int ell_w;
};

// sliced-ELL widths compiled, 3, 5, 7 and 9 (rows are padded to the next
// one; wider rows keep the CSR loop)
constexpr int kEllMax = 9;
inline int ell_width_for(int w) {
  return w <= 3 ? 3 : w <= 5 ? 5 : w <= 7 ? 7 : w <= kEllMax ? kEllMax : 0;
}
constexpr int kBatchRows = 2;

// element j of a per-system base through a 32-bit byte offset (systems
// hold fewer than 2^29 rows)
template <typename T>
__device__ __forceinline__ T ld32(const T* base, int j) {
  return *(const T*)((const char*)base + (uint32_t)j * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st32(T* base, int j, T v) {
  *(T*)((char*)base + (uint32_t)j * (uint32_t)sizeof(T)) = v;
}

template <int BLOCK, int W>
__device__ __forceinline__ void wg_sum(double (&v)[W], double* lds, int& phase) {
  constexpr int NW = BLOCK / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int w = 0; w < W; ++w) v[w] = wave_reduce_sum(v[w]);
  double* buf = lds + phase * NW * W;
  if (lane == 0) {
#pragma unroll
    for (int w = 0; w < W; ++w) buf[wave * W + w] = v[w];
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < W; ++w) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) s += buf[q * W + w];
    v[w] = s;
  }
  phase ^= 1;
}

template <int BLOCK, typename T, int EW, bool OFF16>
__global__ __launch_bounds__(BLOCK) void batch_pcg_kernel(BatchArgs a) {
  constexpr int NW = BLOCK / 64;
  __shared__ double red[2 * NW * 2];
  __shared__ int s_sys;
  extern __shared__ __attribute__((aligned(16))) unsigned char dyn[];
  T* lds_p = (T*)dyn;
  const int tid = threadIdx.x;
  const T* __restrict__ vin = (const T*)a.vin;
  T* __restrict__ R = (T*)a.r;
  T* __restrict__ Dv = (T*)a.d;
  T* __restrict__ P = (T*)a.p;
  T* __restrict__ Q = (T*)a.q;
  const T* __restrict__ invd = (const T*)a.invd;
  int phase = 0;
  for (;;) {
    if (tid == 0) s_sys = atomicAdd(a.counter, 1);
    __syncthreads();
    const int k = s_sys;
    __syncthreads();
    if (k >= a.K) break;
    const int sid = a.order[k];
    const int64_t r0 = a.sys_row[sid];
    const int n = (int)(a.sys_row[sid + 1] - r0);
    const bool in_lds = n <= a.lds_rows;
    // p stored through LDS (small systems) or global memory
    auto pset = [&](int64_t g, T v) {
      if (in_lds) lds_p[g - r0] = v; else P[g] = v;
    };

    // r = b - A x, ||b||, ||r||
    double acc[2] = {0.0, 0.0};
    for (int i = tid; i < n; i += BLOCK) {
      const int64_t g = r0 + i;
      double s = 0.0;
      for (int q = a.rp[g]; q < a.rp[g + 1]; ++q) s += a.v64[q] * a.x[a.ci[q]];
      const double bi = a.b[g], ri = bi - s;
      a.r64[g] = ri;
      acc[0] += bi * bi;
      acc[1] += ri * ri;
    }
    wg_sum<BLOCK, 2>(acc, red, phase);
    const double bnorm = sqrt(acc[0]);
    double rnorm = sqrt(acc[1]);
    int outer = 0, inner_total = 0, conv = 0;
    double rel = bnorm > 0.0 ? rnorm / bnorm : 0.0;
    if (bnorm == 0.0) {
      for (int i = tid; i < n; i += BLOCK) a.x[r0 + i] = 0.0;
      conv = 1;
    } else if (rel < a.tol) {
      conv = 1;
    }
    if (!conv) {
      // Jacobi-PCG in the inner precision with reliable updates: the
      // correction d accumulates in T; when |r| has dropped by inner_tol
      // since the last update (or reaches the tolerance), x += d in FP64,
      // r = b - A x in FP64 replaces the recurrence residual and the
      // search direction is kept, so the Krylov recurrence is not restarted.
      double rz[1] = {0.0};
      for (int i = tid; i < n; i += BLOCK) {
        const int64_t g = r0 + i;
        const T ri = (T)a.r64[g];
        R[g] = ri;
        Dv[g] = (T)0;
        const T zi = invd[g] * ri;
        pset(g, zi);
        rz[0] += (double)ri * (double)zi;
      }
      wg_sum<BLOCK, 1>(rz, red, phase);
      double rzv = rz[0];
      double rupd = rnorm;  // |r| at the last FP64 update
      // The three passes take kBatchRows rows per thread at a time (rows
      // i, i + BLOCK, ...), issuing every load of those rows before the
      // first use; the sums run over the rows in the same order as one row
      // at a time.
      // At 1024 threads (128 VGPRs) one row per lane: two spilled and ran
      // the 4096-system diffusion batch 12 % slower in FP32.
      constexpr int U = BLOCK >= 1024 ? 1 : kBatchRows;
      // per-system bases (uniform) and 32-bit local row offsets: the
      // loads take a scalar base and one 32-bit vector offset
      T* __restrict__ Rs = R + r0;
      T* __restrict__ Ds = Dv + r0;
      T* __restrict__ Ps = P + r0;
      T* __restrict__ Qs = Q + r0;
      const T* __restrict__ Is = invd + r0;
      // the inner loop is compiled twice: p in LDS (ds loads) or in global
      // memory (global loads), never a generic pointer
      auto inner = [&](auto lds_tag) {
        constexpr bool kL = decltype(lds_tag)::value;
        auto pl = [&](int j) -> T {
          if constexpr (kL) return lds_p[j];
          else return ld32((const T*)Ps, j);
        };
        for (int it = 0; it < a.max_inner; ++it) {
          __syncthreads();  // p complete
          double pq[1] = {0.0};
          if constexpr (EW > 0) {
            const int rb = (int)r0;
            const int32_t* __restrict__ cs = a.ell_ci + r0 * EW;
            const int16_t* __restrict__ os = a.ell_off + r0 * EW;
            const T* __restrict__ vs = (const T*)a.ell_v + r0 * EW;
            for (int i = tid; i < n; i += U * BLOCK) {
              int32_t c[U][EW];
              T v[U][EW];
  #pragma unroll
              for (int u = 0; u < U; ++u) {
                int o = i + u * BLOCK;
                if (o < n) {
  #pragma unroll
                  for (int k = 0; k < EW; ++k) {
                    if constexpr (OFF16) c[u][k] = i + u * BLOCK + (int)ld32(os, o);
                    else c[u][k] = ld32(cs, o) - rb;
                    v[u][k] = ld32(vs, o);
                    o += n;
                  }
                }
              }
  #pragma unroll
              for (int u = 0; u < U; ++u) {
                const int j = i + u * BLOCK;
                if (j < n) {
                  T sum = (T)0;
  #pragma unroll
                  for (int k = 0; k < EW; ++k) sum += v[u][k] * pl(c[u][k]);
                  st32(Qs, j, sum);
                  pq[0] += (double)pl(j) * (double)sum;
                }
              }
            }
          } else {
            for (int i = tid; i < n; i += BLOCK) {
              const int64_t g = r0 + i;
              T sum = (T)0;
              for (int q = a.rp[g]; q < a.rp[g + 1]; ++q) sum += vin[q] * pl(a.ci[q] - (int)r0);
              Q[g] = sum;
              pq[0] += (double)pl(i) * (double)sum;
            }
          }
          wg_sum<BLOCK, 1>(pq, red, phase);
          if (!(pq[0] > 0.0)) break;  // breakdown
          const T alpha = (T)(rzv / pq[0]);
          double st[2] = {0.0, 0.0};
          for (int i = tid; i < n; i += U * BLOCK) {
            T dv[U], pv[U], rv[U], qv[U], iv[U];
  #pragma unroll
            for (int u = 0; u < U; ++u) {
              const int j = i + u * BLOCK;
              if (j < n) {
                dv[u] = ld32(Ds, j);
                pv[u] = pl(j);
                rv[u] = ld32(Rs, j);
                qv[u] = ld32(Qs, j);
                iv[u] = ld32(Is, j);
              }
            }
  #pragma unroll
            for (int u = 0; u < U; ++u) {
              const int j = i + u * BLOCK;
              if (j < n) {
                st32(Ds, j, dv[u] + alpha * pv[u]);
                const T ri = rv[u] - alpha * qv[u];
                st32(Rs, j, ri);
                const T zi = iv[u] * ri;
                st[0] += (double)ri * (double)zi;
                st[1] += (double)ri * (double)ri;
              }
            }
          }
          wg_sum<BLOCK, 2>(st, red, phase);
          ++inner_total;
          const double rn = sqrt(st[1]);
          if (rn < a.inner_tol * rupd || rn < a.tol * bnorm) {
            // reliable update: x += d, true residual in FP64
            for (int i = tid; i < n; i += BLOCK) {
              const int64_t g = r0 + i;
              a.x[g] += (double)Dv[g];
              Dv[g] = (T)0;
            }
            __syncthreads();  // x complete
            double rr[1] = {0.0};
            for (int i = tid; i < n; i += BLOCK) {
              const int64_t g = r0 + i;
              double s = 0.0;
              for (int q = a.rp[g]; q < a.rp[g + 1]; ++q) s += a.v64[q] * a.x[a.ci[q]];
              const double ri = a.b[g] - s;
              a.r64[g] = ri;
              rr[0] += ri * ri;
            }
            wg_sum<BLOCK, 1>(rr, red, phase);
            rnorm = sqrt(rr[0]);
            rel = rnorm / bnorm;
            ++outer;
            if (rel < a.tol) {
              conv = 1;
              break;
            }
            if (outer >= a.max_outer) break;
            rupd = rnorm;
            double rzn[1] = {0.0};
            for (int i = tid; i < n; i += BLOCK) {
              const int64_t g = r0 + i;
              const T ri = (T)a.r64[g];
              R[g] = ri;
              rzn[0] += (double)ri * (double)(invd[g] * ri);
            }
            wg_sum<BLOCK, 1>(rzn, red, phase);
            st[0] = rzn[0];
          }
          const T beta = (T)(st[0] / rzv);
          rzv = st[0];
          for (int i = tid; i < n; i += U * BLOCK) {
            T pv[U], rv[U], iv[U];
  #pragma unroll
            for (int u = 0; u < U; ++u) {
              const int j = i + u * BLOCK;
              if (j < n) {
                iv[u] = ld32(Is, j);
                rv[u] = ld32(Rs, j);
                pv[u] = pl(j);
              }
            }
  #pragma unroll
            for (int u = 0; u < U; ++u) {
              const int j = i + u * BLOCK;
              if (j < n) {
                const T pn = iv[u] * rv[u] + beta * pv[u];
                if constexpr (kL) lds_p[j] = pn; else st32(Ps, j, pn);
              }
            }
          }
        }
      };
      if (in_lds) inner(std::true_type{});
      else inner(std::false_type{});
      if (!conv) {  // budget spent: keep the correction made so far
        for (int i = tid; i < n; i += BLOCK) a.x[r0 + i] += (double)Dv[r0 + i];
        __syncthreads();
      }
    }
    if (tid == 0) {
      a.iters[sid] = inner_total;
      a.outer[sid] = outer;
      a.relres[sid] = rel;
      a.conv[sid] = conv;
    }
    __syncthreads();  // the system's LDS p is not reused before all are done
  }
}

__global__ void batch_reset_kernel(int64_t n, double* __restrict__ x,
                                   int* __restrict__ counter) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i == 0) *counter = 0;
  if (i < n) x[i] = 0.0;
}

template <typename T>
__global__ void batch_invd_kernel(int64_t n, const int32_t* __restrict__ rp,
                                  const int32_t* __restrict__ ci,
                                  const double* __restrict__ v, T* __restrict__ out,
                                  int* bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int q = rp[i]; q < rp[i + 1]; ++q)
    if (ci[q] == i) d = v[q];
  if (d == 0.0) atomicOr(bad, 1);
  out[i] = (T)(1.0 / (d == 0.0 ? 1.0 : d));
}

template <typename T>
__global__ void batch_cast_kernel(int64_t n, const double* __restrict__ in,
                                  T* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[i];
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

__global__ void batch_maxlen_kernel(int64_t n, const int32_t* __restrict__ rp,
                                    int* __restrict__ out) {
  int m = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, rp[i + 1] - rp[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// slot k of global row g (local row g - r0 of the system [r0, r1)): the
// row's k-th entry in storage order, or (g, 0); sys_of[g] is found by a
// binary search of the row offsets
template <typename T>
__global__ void batch_ell_kernel(int64_t n, int w, const int64_t* __restrict__ sys_row,
                                 int K, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci,
                                 const double* __restrict__ v,
                                 int32_t* __restrict__ eci, int16_t* __restrict__ eoff,
                                 int* off_bad, T* __restrict__ ev) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = K;  // largest k with sys_row[k] <= g
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sys_row[mid] <= g) lo = mid; else hi = mid;
    }
    const int64_t r0 = sys_row[lo], len_sys = sys_row[lo + 1] - r0;
    const int b = rp[g], len = rp[g + 1] - b;
    for (int k = 0; k < w; ++k) {
      const int64_t o = r0 * w + (int64_t)k * len_sys + (g - r0);
      if (k < len) {
        if (eci) eci[o] = ci[b + k];
        if (eoff) {
          const int64_t d = (int64_t)ci[b + k] - g;
          if (d < -32768 || d > 32767) atomicOr(off_bad, 1);
          else eoff[o] = (int16_t)d;
        }
        ev[o] = (T)v[b + k];
      } else {
        if (eci) eci[o] = (int32_t)g;
        if (eoff) eoff[o] = 0;
        ev[o] = (T)0;
      }
    }
  }
}

}  // namespace
}  // namespace spk

using namespace spk;

struct spk_batch_s {
  spk_csr_s* A = nullptr;  // owned
  int64_t K = 0, total = 0;
  std::vector<int64_t> h_row;  // K + 1
  DevBuf<int64_t> sys_row;
  DevBuf<int32_t> order;       // by decreasing size
  std::vector<int32_t> h_order;
  DevBuf<double> r64;
  DevBuf<float> r32, d32, p32, q32, invd32;
  DevBuf<double> rI, dI, pI, qI, invd64;  // FP64 inner vectors, made on first use
  DevBuf<int> counter;
  // sliced-ELL copy of the matrix for the inner loop (rows of at most
  // kEllMax entries); the FP64 values are made on first use
  int ell_w = 0;
  DevBuf<int32_t> ell_ci;
  DevBuf<int16_t> ell_off;  // when every column is within 32767 of its row
  DevBuf<float> ell_v32;
  DevBuf<double> ell_v64;
  DevBuf<int32_t> iters, outer, conv;
  DevBuf<double> relres;
  spk_batch_tuning tuning{256, 2, 0, 0};
  // captured graphs, keyed by (b, x, config, tuning, precision)
  struct Key {
    const void* b;
    void* x;
    double tol, itol;
    int64_t mo, mi;
    int block, wg, lds, prec;
    bool operator<(const Key& o) const {
      return std::tie(b, x, tol, itol, mo, mi, block, wg, lds, prec) <
             std::tie(o.b, o.x, o.tol, o.itol, o.mo, o.mi, o.block, o.wg, o.lds,
                      o.prec);
    }
  };
  std::map<Key, hipGraphExec_t> graphs;

  ~spk_batch_s() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    delete A;
  }

  void setup() {
    hipStream_t s = stream();
    const size_t n = (size_t)total;
    // the inner loop addresses a system's vectors with 32-bit byte offsets
    int64_t max_rows = 0;
    for (int64_t k = 0; k < K; ++k)
      max_rows = std::max(max_rows, h_row[(size_t)k + 1] - h_row[(size_t)k]);
    SPK_REQUIRE(max_rows < ((int64_t)1 << 29), SPK_INVALID_ARGUMENT,
                "a system of the batch holds 2^29 rows or more");
    sys_row.alloc((size_t)K + 1);
    SPK_HIP_TRY(hipMemcpyAsync(sys_row.p, h_row.data(), (K + 1) * 8,
                               hipMemcpyHostToDevice, s));
    h_order.resize((size_t)K);
    for (int64_t k = 0; k < K; ++k) h_order[(size_t)k] = (int32_t)k;
    std::stable_sort(h_order.begin(), h_order.end(), [&](int32_t a, int32_t b) {
      return h_row[a + 1] - h_row[a] > h_row[b + 1] - h_row[b];
    });
    order.alloc((size_t)K);
    SPK_HIP_TRY(hipMemcpyAsync(order.p, h_order.data(), K * 4,
                               hipMemcpyHostToDevice, s));
    r64.alloc(n);
    r32.alloc(n);
    d32.alloc(n);
    p32.alloc(n);
    q32.alloc(n);
    invd32.alloc(n);
    invd64.alloc(n);
    counter.alloc(1);
    iters.alloc((size_t)K);
    outer.alloc((size_t)K);
    conv.alloc((size_t)K);
    relres.alloc((size_t)K);
    DevBuf<int> bad(1);
    SPK_HIP_TRY(hipMemsetAsync(bad.p, 0, 4, s));
    hipLaunchKernelGGL(batch_invd_kernel<double>, dim3(nblk(total)), dim3(256), 0,
                       s, total, (const int32_t*)A->rp.p, (const int32_t*)A->ci.p,
                       (const double*)A->v64.p, invd64.p, bad.p);
    hipLaunchKernelGGL(batch_cast_kernel<float>, dim3(nblk(total)), dim3(256), 0,
                       s, total, (const double*)invd64.p, invd32.p);
    SPK_HIP_TRY(hipGetLastError());
    int hb = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&hb, bad.p, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    SPK_REQUIRE(!hb, SPK_INVALID_ARGUMENT, "zero diagonal element in Jacobi");
    (void)A->vals32();
    // the sliced-ELL copy when every row is short
    DevBuf<int> mx(1);
    SPK_HIP_TRY(hipMemsetAsync(mx.p, 0, 4, s));
    hipLaunchKernelGGL(batch_maxlen_kernel, dim3(std::min<unsigned>(nblk(total), 4096)),
                       dim3(256), 0, s, total, (const int32_t*)A->rp.p, mx.p);
    SPK_HIP_TRY(hipGetLastError());
    int w = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&w, mx.p, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    // ... and a system's sliced-ELL slots likewise (FP64 values: 8 bytes)
    if (w >= 1 && ell_width_for(w) > 0 &&
        (int64_t)total * ell_width_for(w) < ((int64_t)1 << 31) &&
        max_rows * ell_width_for(w) * 8 < ((int64_t)1 << 32)) {
      w = ell_width_for(w);
      ell_w = w;
      ell_ci.alloc(n * (size_t)w);
      ell_off.alloc(n * (size_t)w);
      ell_v32.alloc(n * (size_t)w);
      DevBuf<int> obad(1);
      SPK_HIP_TRY(hipMemsetAsync(obad.p, 0, 4, s));
      hipLaunchKernelGGL(batch_ell_kernel<float>, dim3(std::min<unsigned>(nblk(total), 16384)),
                         dim3(256), 0, s, total, w, (const int64_t*)sys_row.p, (int)K,
                         (const int32_t*)A->rp.p,
                         (const int32_t*)A->ci.p, (const double*)A->v64.p, ell_ci.p,
                         ell_off.p, obad.p, ell_v32.p);
      SPK_HIP_TRY(hipGetLastError());
      int ob = 0;
      SPK_HIP_TRY(hipMemcpyAsync(&ob, obad.p, 4, hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      // 16-bit offsets when they all fit (2 bytes per slot instead of 4)
      if (ob) ell_off.release();
      else ell_ci.release();
    }
  }

  void ensure_f64() {
    if (!rI.n) {
      const size_t n = (size_t)total;
      rI.alloc(n);
      dI.alloc(n);
      pI.alloc(n);
      qI.alloc(n);
    }
    if (ell_w > 0 && !ell_v64.n) {
      ell_v64.alloc((size_t)total * (size_t)ell_w);
      hipLaunchKernelGGL(batch_ell_kernel<double>,
                         dim3(std::min<unsigned>(nblk(total), 16384)), dim3(256), 0,
                         stream(), total, ell_w, (const int64_t*)sys_row.p, (int)K,
                         (const int32_t*)A->rp.p,
                         (const int32_t*)A->ci.p, (const double*)A->v64.p,
                         (int32_t*)nullptr, (int16_t*)nullptr, (int*)nullptr,
                         ell_v64.p);
      SPK_HIP_TRY(hipGetLastError());
    }
  }

  BatchArgs args(const spk_batch_config& c, const double* b, double* x,
                 int prec, const int32_t* ord, int K_) {
    BatchArgs a{};
    a.rp = A->rp.p;
    a.ci = A->ci.p;
    a.v64 = A->v64.p;
    a.sys_row = sys_row.p;
    a.order = ord;
    a.K = K_;
    a.b = b;
    a.x = x;
    a.r64 = r64.p;
    if (prec == SPK_F32) {
      a.vin = A->vals32();
      a.r = r32.p;
      a.d = d32.p;
      a.p = p32.p;
      a.q = q32.p;
      a.invd = invd32.p;
    } else {
      ensure_f64();
      a.vin = A->v64.p;
      a.r = rI.p;
      a.d = dI.p;
      a.p = pI.p;
      a.q = qI.p;
      a.invd = invd64.p;
    }
    a.ell_w = ell_w;
    a.ell_n = total;
    a.ell_ci = ell_ci.p;
    a.ell_off = ell_off.p;
    a.ell_v = prec == SPK_F32 ? (const void*)ell_v32.p : (const void*)ell_v64.p;
    a.counter = counter.p;
    a.iters = iters.p;
    a.outer = outer.p;
    a.relres = relres.p;
    a.conv = conv.p;
    a.tol = c.tolerance;
    a.inner_tol = c.inner_tolerance;
    a.max_outer = (int)c.max_outer;
    a.max_inner = (int)c.max_inner;
    return a;
  }

  // enqueue reset + persistent kernel on the library stream
  void enqueue(const spk_batch_config& c, const double* b, double* x,
               const spk_batch_tuning& t, const int32_t* ord, int K_,
               int64_t reset_rows) {
    hipStream_t s = stream();
    const int prec = c.precision;
    BatchArgs a = args(c, b, x, prec, ord, K_);
    // the LDS bin threshold is capped to the budget of this precision
    const size_t es = prec == SPK_F32 ? 4 : 8;
    const int64_t cap = 131072 / ((int64_t)es * std::max(1, t.wg_per_cu));
    a.lds_rows = (int)std::min<int64_t>(t.lds_rows, cap);
    const size_t dyn = (size_t)a.lds_rows * es;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(nblk(std::max<int64_t>(reset_rows, 1))),
                       dim3(256), 0, s, reset_rows, x, counter.p);
    const int grid = std::max(1, t.wg_per_cu * ctx().num_cus);
#define SPK_BATCH_EW(BS, TT, EW)                                                    \
  do {                                                                              \
    if constexpr (EW > 0) {                                                         \
      if (a.ell_off)                                                                \
        hipLaunchKernelGGL((batch_pcg_kernel<BS, TT, EW, true>), dim3(grid),        \
                           dim3(BS), dyn, s, a);                                    \
      else                                                                          \
        hipLaunchKernelGGL((batch_pcg_kernel<BS, TT, EW, false>), dim3(grid),       \
                           dim3(BS), dyn, s, a);                                    \
    } else {                                                                        \
      hipLaunchKernelGGL((batch_pcg_kernel<BS, TT, 0, false>), dim3(grid), dim3(BS), \
                         dyn, s, a);                                                \
    }                                                                               \
  } while (0)
#define SPK_BATCH_LAUNCH(BS, TT)                     \
  do {                                               \
    switch (a.ell_w) {                               \
      case 3: SPK_BATCH_EW(BS, TT, 3); break;        \
      case 5: SPK_BATCH_EW(BS, TT, 5); break;        \
      case 7: SPK_BATCH_EW(BS, TT, 7); break;        \
      case 9: SPK_BATCH_EW(BS, TT, 9); break;        \
      default: SPK_BATCH_EW(BS, TT, 0); break;       \
    }                                                \
  } while (0)
    if (prec == SPK_F32) {
      if (t.block == 1024) SPK_BATCH_LAUNCH(1024, float);
      else if (t.block == 512) SPK_BATCH_LAUNCH(512, float);
      else SPK_BATCH_LAUNCH(256, float);
    } else {
      if (t.block == 1024) SPK_BATCH_LAUNCH(1024, double);
      else if (t.block == 512) SPK_BATCH_LAUNCH(512, double);
      else SPK_BATCH_LAUNCH(256, double);
    }
#undef SPK_BATCH_LAUNCH
#undef SPK_BATCH_EW
    SPK_HIP_TRY(hipGetLastError());
  }
};

namespace {

void check_tuning(const spk_batch_tuning& t, int prec) {
  SPK_REQUIRE(t.block == 256 || t.block == 512 || t.block == 1024,
              SPK_INVALID_ARGUMENT, "block must be 256, 512 or 1024");
  SPK_REQUIRE(t.wg_per_cu >= 1 && t.wg_per_cu * t.block <= 2048,
              SPK_INVALID_ARGUMENT, "at most 2048 threads per CU");
  const int64_t es = prec == SPK_F32 ? 4 : 8;
  SPK_REQUIRE(t.lds_rows >= 0 && (int64_t)t.lds_rows * es * t.wg_per_cu <= 131072,
              SPK_INVALID_ARGUMENT, "lds_rows exceeds the LDS budget");
}

void check_config(const spk_batch_config* c) {
  SPK_REQUIRE(c, SPK_INVALID_ARGUMENT, "null configuration");
  SPK_REQUIRE(c->max_outer >= 0 && c->max_inner >= 0 && c->max_outer < INT32_MAX &&
                  c->max_inner < INT32_MAX,
              SPK_INVALID_ARGUMENT, "bad iteration budget");
  SPK_REQUIRE(c->precision == SPK_F32 || c->precision == SPK_F64,
              SPK_INVALID_ARGUMENT, "bad precision");
}

spk_batch_s* make_batch(spk_csr_s* A, std::vector<int64_t> rows) {
  auto* B = new spk_batch_s();
  B->A = A;
  B->K = (int64_t)rows.size() - 1;
  B->total = rows.back();
  B->h_row = std::move(rows);
  try {
    B->setup();
  } catch (...) {
    B->A = nullptr;
    delete B;
    throw;
  }
  return B;
}

}  // namespace

extern "C" {

spk_status spk_batch_create_diffusion(int64_t count, uint64_t seed, int64_t side_min,
                                      int64_t side_max, spk_batch* out) {
  return guarded([&] {
    SPK_REQUIRE(out && count > 0 && count < INT32_MAX && side_min >= 2 &&
                    side_max >= side_min && side_max <= 46340,
                SPK_INVALID_ARGUMENT, "bad argument");
    std::vector<int32_t> side((size_t)count);
    std::vector<int64_t> row((size_t)count + 1), nz((size_t)count + 1);
    row[0] = nz[0] = 0;
    for (int64_t k = 0; k < count; ++k) {
      const int64_t s = side_min + (int64_t)(hmix(seed ^ 0x534944ULL, (uint64_t)k) %
                                             (uint64_t)(side_max - side_min + 1));
      side[(size_t)k] = (int32_t)s;
      row[(size_t)k + 1] = row[(size_t)k] + s * s;
      nz[(size_t)k + 1] = nz[(size_t)k] + 5 * s * s - 4 * s;
    }
    const int64_t total = row.back(), nnz = nz.back();
    SPK_REQUIRE(total < INT32_MAX && nnz < INT32_MAX, SPK_INVALID_ARGUMENT,
                "batch exceeds int32 indexing");
    // conductivities on the host, threads over systems
    std::vector<double> kap((size_t)total);
    {
      unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      std::vector<std::thread> th;
      for (unsigned w = 0; w < nt; ++w)
        th.emplace_back([&, w] {
          for (int64_t k = w; k < count; k += nt) {
            const uint64_t t = hmix(seed ^ 0x424154ULL, (uint64_t)k);
            const int64_t s = side[(size_t)k];
            double* o = kap.data() + row[(size_t)k];
            for (int64_t c = 0; c < s * s; ++c) {
              const double u1 = 1.0 - hu01(hmix(t, 2 * (uint64_t)c));
              const double u2 = hu01(hmix(t, 2 * (uint64_t)c + 1));
              const double z =
                  std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
              o[c] = std::exp(z);
            }
          }
        });
      for (auto& t : th) t.join();
    }
    hipStream_t s = stream();
    auto* A = new spk_csr_s();
    spk_batch_s* B = nullptr;
    try {
      A->allocate(total, total, nnz);
      DevBuf<double> dk((size_t)total);
      DevBuf<int64_t> drow((size_t)count + 1), dnz((size_t)count + 1);
      DevBuf<int32_t> dside((size_t)count);
      SPK_HIP_TRY(hipMemcpyAsync(dk.p, kap.data(), total * 8, hipMemcpyHostToDevice, s));
      SPK_HIP_TRY(hipMemcpyAsync(drow.p, row.data(), (count + 1) * 8,
                                 hipMemcpyHostToDevice, s));
      SPK_HIP_TRY(hipMemcpyAsync(dnz.p, nz.data(), (count + 1) * 8,
                                 hipMemcpyHostToDevice, s));
      SPK_HIP_TRY(hipMemcpyAsync(dside.p, side.data(), count * 4,
                                 hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(bd_fill_kernel, dim3(nblk(total)), dim3(256), 0, s, total,
                         (int)count, (const int64_t*)drow.p, (const int64_t*)dnz.p,
                         (const int32_t*)dside.p, (const double*)dk.p, A->rp.p,
                         A->ci.p, A->v64.p);
      SPK_HIP_TRY(hipGetLastError());
      SPK_HIP_TRY(hipStreamSynchronize(s));
      finish_store(*A, false);
      B = make_batch(A, std::move(row));
    } catch (...) {
      delete A;
      throw;
    }
    *out = B;
  });
}

spk_status spk_batch_create(spk_csr A, int64_t count, const int64_t* row_offsets,
                            spk_batch* out) {
  return guarded([&] {
    SPK_REQUIRE(A && out && row_offsets && count > 0 && count < INT32_MAX,
                SPK_INVALID_ARGUMENT, "bad argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "batch matrix must be square");
    std::vector<int64_t> row(row_offsets, row_offsets + count + 1);
    SPK_REQUIRE(row[0] == 0 && row.back() == A->nrows, SPK_INVALID_ARGUMENT,
                "row offsets must run from 0 to the number of rows");
    for (int64_t k = 0; k < count; ++k)
      SPK_REQUIRE(row[(size_t)k + 1] > row[(size_t)k], SPK_INVALID_ARGUMENT,
                  "every system needs at least one row");
    // copy the store so the batch owns it
    spk_csr_s* C = nullptr;
    spk_status st = spk_csr_create_device(A->nrows, A->ncols, A->nnz, A->rp.p, A->ci.p,
                                          A->v64.p, &C);
    if (st != SPK_OK) throw Error{st, spk_last_error()};
    {
      // block check on the host copy of the pattern
      std::vector<int32_t> hci((size_t)C->nnz);
      if (C->nnz)
        SPK_HIP_TRY(hipMemcpy(hci.data(), C->ci.p, C->nnz * 4, hipMemcpyDeviceToHost));
      bool ok = true;
      for (int64_t k = 0; k < count && ok; ++k)
        for (int64_t i = row[(size_t)k]; i < row[(size_t)k + 1] && ok; ++i)
          for (int32_t q = C->h_rp[(size_t)i]; q < C->h_rp[(size_t)i + 1]; ++q)
            if (hci[(size_t)q] < row[(size_t)k] || hci[(size_t)q] >= row[(size_t)k + 1]) {
              ok = false;
              break;
            }
      if (!ok) {
        delete C;
        SPK_REQUIRE(false, SPK_INVALID_ARGUMENT,
                    "an entry couples two systems (the matrix must be block diagonal)");
      }
    }
    *out = make_batch(C, std::move(row));
  });
}

spk_status spk_batch_destroy(spk_batch B) {
  return guarded([&] {
    if (!B) return;
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    delete B;
  });
}

spk_status spk_batch_info(spk_batch B, int64_t* count, int64_t* total_rows,
                          int64_t* nnz, spk_csr* matrix) {
  return guarded([&] {
    SPK_REQUIRE(B, SPK_INVALID_ARGUMENT, "null argument");
    if (count) *count = B->K;
    if (total_rows) *total_rows = B->total;
    if (nnz) *nnz = B->A->nnz;
    if (matrix) *matrix = B->A;
  });
}

spk_status spk_batch_row_offsets(spk_batch B, int64_t* row_offsets) {
  return guarded([&] {
    SPK_REQUIRE(B && row_offsets, SPK_INVALID_ARGUMENT, "null argument");
    std::copy(B->h_row.begin(), B->h_row.end(), row_offsets);
  });
}

spk_status spk_batch_set_tuning(spk_batch B, const spk_batch_tuning* t) {
  return guarded([&] {
    SPK_REQUIRE(B && t, SPK_INVALID_ARGUMENT, "null argument");
    check_tuning(*t, SPK_F32);
    B->tuning = *t;
  });
}

spk_status spk_batch_get_tuning(spk_batch B, spk_batch_tuning* t) {
  return guarded([&] {
    SPK_REQUIRE(B && t, SPK_INVALID_ARGUMENT, "null argument");
    *t = B->tuning;
  });
}

spk_status spk_batch_pcg_solve(spk_batch B, const spk_batch_config* cfg,
                               const double* d_b, double* d_x, int use_graph,
                               spk_batch_summary* out) {
  return guarded([&] {
    TraceRange trace_("spk_batch_pcg_solve");
    SPK_REQUIRE(B && d_b && d_x, SPK_INVALID_ARGUMENT, "null argument");
    check_config(cfg);
    check_tuning(B->tuning, SPK_F32);  // lds_rows is capped per precision
    if (cfg->precision == SPK_F64) B->ensure_f64();  // no allocation in a capture
    hipStream_t s = stream();
    const spk_batch_tuning t = B->tuning;
    if (use_graph) {
      spk_batch_s::Key key{d_b, d_x, cfg->tolerance, cfg->inner_tolerance,
                           cfg->max_outer, cfg->max_inner, t.block, t.wg_per_cu,
                           t.lds_rows, cfg->precision};
      auto it = B->graphs.find(key);
      if (it == B->graphs.end()) {
        TraceRange capture_("batch: graph capture");
        hipGraph_t g = nullptr;
        SPK_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          B->enqueue(*cfg, d_b, d_x, t, B->order.p, (int)B->K, B->total);
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        SPK_HIP_TRY(hipStreamEndCapture(s, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SPK_HIP_TRY(e);
        it = B->graphs.emplace(key, ex).first;
      }
      TraceRange replay_("batch: graph replay");
      SPK_HIP_TRY(hipGraphLaunch(it->second, s));
    } else {
      B->enqueue(*cfg, d_b, d_x, t, B->order.p, (int)B->K, B->total);
    }
    if (out) {
      std::vector<int32_t> it((size_t)B->K), ou((size_t)B->K), cv((size_t)B->K);
      std::vector<double> rr((size_t)B->K);
      SPK_HIP_TRY(hipMemcpyAsync(it.data(), B->iters.p, B->K * 4, hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipMemcpyAsync(ou.data(), B->outer.p, B->K * 4, hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipMemcpyAsync(cv.data(), B->conv.p, B->K * 4, hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipMemcpyAsync(rr.data(), B->relres.p, B->K * 8, hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      *out = spk_batch_summary{};
      for (int64_t k = 0; k < B->K; ++k) {
        out->total_inner_iterations += it[(size_t)k];
        out->max_inner_iterations = std::max<int64_t>(out->max_inner_iterations, it[(size_t)k]);
        out->total_outer_iterations += ou[(size_t)k];
        out->converged += cv[(size_t)k];
        out->max_relres = std::max(out->max_relres, rr[(size_t)k]);
      }
    }
  });
}

spk_status spk_batch_results(spk_batch B, int32_t* inner_iterations,
                             int32_t* outer_iterations, double* relres,
                             int32_t* converged) {
  return guarded([&] {
    SPK_REQUIRE(B, SPK_INVALID_ARGUMENT, "null argument");
    hipStream_t s = stream();
    if (inner_iterations)
      SPK_HIP_TRY(hipMemcpyAsync(inner_iterations, B->iters.p, B->K * 4,
                                 hipMemcpyDeviceToHost, s));
    if (outer_iterations)
      SPK_HIP_TRY(hipMemcpyAsync(outer_iterations, B->outer.p, B->K * 4,
                                 hipMemcpyDeviceToHost, s));
    if (relres)
      SPK_HIP_TRY(hipMemcpyAsync(relres, B->relres.p, B->K * 8, hipMemcpyDeviceToHost, s));
    if (converged)
      SPK_HIP_TRY(hipMemcpyAsync(converged, B->conv.p, B->K * 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
  });
}

spk_status spk_batch_autotune(spk_batch B, const spk_batch_config* cfg,
                              const double* d_b, double* d_x, int64_t sample,
                              spk_batch_tuning* chosen, spk_batch_tuning* tried,
                              float* tried_ms, int64_t* ntried) {
  return guarded([&] {
    TraceRange trace_("spk_batch_autotune");
    SPK_REQUIRE(B && d_b && d_x && chosen, SPK_INVALID_ARGUMENT, "null argument");
    check_config(cfg);
    if (cfg->precision == SPK_F64) B->ensure_f64();
    const int64_t m = std::max<int64_t>(1, std::min<int64_t>(sample, B->K));
    // stratified sample of the size-ordered systems
    std::vector<int32_t> ord((size_t)m);
    for (int64_t j = 0; j < m; ++j)
      ord[(size_t)j] = B->h_order[(size_t)((j * B->K) / m)];
    DevBuf<int32_t> dord((size_t)m);
    hipStream_t s = stream();
    SPK_HIP_TRY(hipMemcpyAsync(dord.p, ord.data(), m * 4, hipMemcpyHostToDevice, s));
    std::vector<spk_batch_tuning> cands;
    for (int bs : {256, 512, 1024})
      for (int wg : {1, 2, 4})
        for (int lds : {0, 16384}) {
          spk_batch_tuning t{bs, wg, lds, 0};
          if (bs * wg > 2048) continue;
          const int64_t es = cfg->precision == SPK_F32 ? 4 : 8;
          if ((int64_t)lds * es * wg > 131072) continue;
          cands.push_back(t);
        }
    hipEvent_t e0, e1;
    SPK_HIP_TRY(hipEventCreate(&e0));
    SPK_HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    spk_batch_tuning bt = cands[0];
    int64_t k = 0;
    try {
      // one untimed run, then one timed run per candidate (x is reset by
      // the enqueued reset kernel over all rows)
      B->enqueue(*cfg, d_b, d_x, cands[0], dord.p, (int)m, B->total);
      for (const auto& t : cands) {
        SPK_HIP_TRY(hipEventRecord(e0, s));
        B->enqueue(*cfg, d_b, d_x, t, dord.p, (int)m, B->total);
        SPK_HIP_TRY(hipEventRecord(e1, s));
        SPK_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        SPK_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (tried && tried_ms) {
          tried[k] = t;
          tried_ms[k] = ms;
        }
        ++k;
        if (ms < best) {
          best = ms;
          bt = t;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ntried) *ntried = k;
    B->tuning = bt;
    *chosen = bt;
  });
}

}  // extern "C"
