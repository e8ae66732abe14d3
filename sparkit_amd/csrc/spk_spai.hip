// spk_spai.hip -- learned sparse approximate inverse: a small MLP predicts
// the rows of M on A's sparsity pattern from per-row features; M is then
// applied as one SpMV (a spk_precond of kind "spmv").  The network is trained
// on the device with hand-written BF16 MFMA forward/backward kernels and an
// Adam update, minimising (1/n) ||I - M A||_F^2.
//
// Row slots.  Row i must hold its diagonal, a_ii > 0, and at most 4 entries
// on each side of it.  Slot 4 is the diagonal; the entries left of it, from
// the nearest outwards, take slots 3, 2, 1, 0; those right of it 5..8 (for
// an interior row of a 9-point stencil: SW S SE W C E NW N NE).
// Features (24, computed in double, stored as float) for the off-diagonal
// slots k = 0..7 (slots 0 1 2 3 5 6 7 8) with column j, zero when empty:
//   x[k] = a_ij / a_ii,  x[8+k] = (a_jj - a_ii) / (a_jj + a_ii),
//   x[16+k] = rowsum_j / a_jj   (rowsum in storage order).
// Network: y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3, 24 -> 64 -> 64
// -> 9; GEMM operands (features, weights, activations, backward deltas) are
// bfloat16 (round to nearest even), accumulation and bias/elementwise work
// fp32.  M_i,slot = (delta_{slot,4} + y_slot) / a_ii: zero output weights
// give the Jacobi preconditioner.
//
// Kernel layout.  One workgroup = 4 wavefronts; each wavefront owns a tile
// of 64 rows at a time (lane = row for the per-row work) and a 32 KB LDS
// region for the tile's operands; the bf16 weights sit in LDS once per
// workgroup.  Every GEMM is a sequence of v_mfma_f32_16x16x32_bf16 on 16x16
// output tiles whose A/B fragments are read from LDS:
//   forward   H1 = relu(X W1 + b1), H2 = relu(H1 W2 + b2), Y = H2 W3 + b3
//   per row   m from Y, (M A)_i by a merge of the slot rows, dL/dy
//   backward  dW3 += H2^T dY; dH2 = dY W3^T * [H2 > 0]; dW2 += H1^T dH2;
//             dH1 = dH2 W2^T * [H1 > 0]; dW1 += X^T dH1
// Weight gradients stay in the MFMA accumulators across a wavefront's
// tiles, are summed over the workgroup's wavefronts in LDS in a fixed order
// and written as one partial per workgroup; the update kernel sums the
// partials in workgroup order (deterministic) and applies Adam to the fp32
// master weights, then refreshes the bf16 copy.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "spk_csr.hpp"
#include "spk_vec.hpp"

namespace spk {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kF = 24, kH = 64, kO = 9;
// canonical parameter layout: W1[24][64] b1 W2[64][64] b2 W3[64][9] b3
constexpr int kW1 = 0, kB1 = kW1 + kF * kH, kW2 = kB1 + kH, kB2 = kW2 + kH * kH,
              kW3 = kB2 + kH, kB3 = kW3 + kH * kO, kNP = kB3 + kO;
static_assert(kNP == SPK_SPAI_NPARAM, "parameter count");
// padded device shapes
constexpr int kFp = 32;  // X columns / W1 rows
constexpr int kOk = 32;  // dY columns / W3 columns (K of dY W3^T)
constexpr int kOn = 16;  // Y columns (one output tile)
// gradient partial layout (floats)
constexpr int gW1 = 0, gW2 = gW1 + kFp * kH, gW3 = gW2 + kH * kH,
              gB1 = gW3 + kH * kOn, gB2 = gB1 + kH, gB3 = gB2 + kH,
              kGradLen = gB3 + kOn;
constexpr int kTile = 64;   // rows per wavefront tile
constexpr int kWaves = 4;   // wavefronts per workgroup
// per-wavefront LDS region, in bf16 elements
constexpr int oX = 0, oH1 = oX + kTile * kFp, oH2 = oH1 + kTile * kH,
              oYD = oH2 + kTile * kH, oD = oYD + kTile * kOk,
              kPool = oD + kTile * kH;

struct Packed {  // what the tile kernel reads: bf16 weights, fp32 biases
  bf16 w1[kFp * kH];
  bf16 w2[kH * kH];
  bf16 w3[kH * kOk];
  float b1[kH], b2[kH], b3[kOn];
};

__host__ __device__ inline uint16_t bf16_bits(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)(u >> 16);  // NaN
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ bf16 to_bf16(float x) {
  return __builtin_bit_cast(bf16, bf16_bits(x));
}
__device__ __forceinline__ float from_bf16(bf16 h) {
  return __builtin_bit_cast(float, (uint32_t)__builtin_bit_cast(uint16_t, h)
                                       << 16);
}

// acc += A[16 x 32] B[32 x 16]; A(r, k) = a[r*ars + k*aks],
// B(k, c) = b[k*bks + c*bcs].  Lane l supplies row/column l % 16 and the
// eight k of group l / 16; acc[i] is C(4 (l / 16) + i, l % 16).
__device__ __forceinline__ floatx4 mma(const bf16* a, int ars, int aks,
                                       const bf16* b, int bks, int bcs,
                                       floatx4 acc) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  bf16x8 av, bv;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    av[j] = a[r * ars + (8 * g + j) * aks];
    bv[j] = b[(8 * g + j) * bks + r * bcs];
  }
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
}

__device__ __forceinline__ floatx4 splat(float v) { return floatx4{v, v, v, v}; }

// per-matrix pass: diagonal, row sums, slot-rule check
__global__ void spai_prepare_kernel(int64_t n, const int32_t* __restrict__ rp,
                                    const int32_t* __restrict__ ci,
                                    const double* __restrict__ v,
                                    double* __restrict__ diag,
                                    double* __restrict__ rsum, int* err) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p0 = rp[i], p1 = rp[i + 1];
  int dpos = -1;
  double s = 0.0;
  for (int p = p0; p < p1; ++p) {
    if (ci[p] == i) dpos = p;
    s += v[p];
  }
  const bool bad = dpos < 0 || !(v[dpos] > 0.0) || dpos - p0 > 4 ||
                   p1 - dpos - 1 > 4;
  if (bad) atomicOr(err, 1);
  diag[i] = bad ? 1.0 : v[dpos];
  rsum[i] = s;
}

struct TileArgs {
  int64_t n;
  const int32_t* __restrict__ rp;
  const int32_t* __restrict__ ci;
  const double* __restrict__ v;
  const double* __restrict__ diag;
  const double* __restrict__ rsum;
  const Packed* __restrict__ pk;
  double* __restrict__ mvals;    // predict: M values on A's pattern
  float* __restrict__ partials;  // train: [gridDim.x][kGradLen]
  double* __restrict__ loss_part;  // train: [gridDim.x]
  double scale;                  // train: 1 / rows of the batch
  int accumulate;                // train: add to the existing partials
  // train, symmetric objective: dL/dm per entry of A's pattern and each
  // row's loss, computed beforehand (null: the row merge below)
  const double* __restrict__ gm;
  const double* __restrict__ rowloss;
};

// Row i's residual e_i - sum_s m_s A_{j_s,:} by a merge of the slot rows
// (their columns ascend): returns ||.||^2 and adds dL/dm_s to gs[s].
__device__ __forceinline__ double row_residual(const int32_t* __restrict__ rp,
                                               const int32_t* __restrict__ ci,
                                               const double* __restrict__ v,
                                               int64_t i, const int* pos,
                                               const double* m, double* gs) {
#pragma clang fp contract(off)
  int hp[kO], he[kO], hc[kO];
  double hv[kO];
#pragma unroll
  for (int s = 0; s < kO; ++s) {
    hp[s] = he[s] = 0;
    if (pos[s] >= 0) {
      const int j = ci[pos[s]];
      hp[s] = rp[j];
      he[s] = rp[j + 1];
    }
    hc[s] = hp[s] < he[s] ? ci[hp[s]] : INT32_MAX;
    hv[s] = hp[s] < he[s] ? v[hp[s]] : 0.0;
  }
  double lrow = 0.0;
  for (;;) {
    int c = INT32_MAX;
#pragma unroll
    for (int s = 0; s < kO; ++s) c = hc[s] < c ? hc[s] : c;
    if (c == INT32_MAX) break;
    double tsum = 0.0;
#pragma unroll
    for (int s = 0; s < kO; ++s)
      if (hc[s] == c) tsum += m[s] * hv[s];
    const double res = ((int64_t)c == i ? 1.0 : 0.0) - tsum;
    lrow += res * res;
#pragma unroll
    for (int s = 0; s < kO; ++s)
      if (hc[s] == c) {
        gs[s] += -2.0 * res * hv[s];
        ++hp[s];
        hc[s] = hp[s] < he[s] ? ci[hp[s]] : INT32_MAX;
        hv[s] = hp[s] < he[s] ? v[hp[s]] : 0.0;
      }
  }
  return lrow;
}

// Row i's slot positions: pos[4] its diagonal, pos[3..0] the entries left
// of it from the nearest outwards, pos[5..8] those right of it (-1 unused).
__device__ __forceinline__ void row_slots(const int32_t* __restrict__ rp,
                                          const int32_t* __restrict__ ci, int64_t i,
                                          int* pos) {
  const int p0 = rp[i], p1 = rp[i + 1];
  int d = p0;
  while (ci[d] != (int32_t)i) ++d;  // present: checked by the prepare pass
  pos[4] = d;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    pos[3 - k] = (d - 1 - k >= p0) ? d - 1 - k : -1;
    pos[5 + k] = (d + 1 + k < p1) ? d + 1 + k : -1;
  }
}

template <bool TRAIN>
__global__ __launch_bounds__(256) void spai_tile_kernel(TileArgs A) {
#pragma clang fp contract(off)
  __shared__ Packed sw;
  __shared__ __attribute__((aligned(16))) bf16 pool[kWaves * kPool];
  __shared__ float b3tmp[kTile * kO];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  {  // weights into LDS
    const uint32_t* src = (const uint32_t*)A.pk;
    uint32_t* dst = (uint32_t*)&sw;
    for (int k = threadIdx.x; k < (int)(sizeof(Packed) / 4); k += blockDim.x)
      dst[k] = src[k];
  }
  __syncthreads();
  bf16* X = pool + wave * kPool + oX;
  bf16* H1 = pool + wave * kPool + oH1;
  bf16* H2 = pool + wave * kPool + oH2;
  bf16* dY = pool + wave * kPool + oYD;
  float* Y = (float*)dY;  // Y[64][16] fp32 shares the bytes of dY[64][32]
  bf16* D = pool + wave * kPool + oD;

  floatx4 aW1[2][4], aW2[4][4], aW3[4];
  float ab1[4], ab2[4], ab3[kO];
  double lsum = 0.0;
  if (TRAIN) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      aW3[a] = splat(0.f);
      ab1[a] = ab2[a] = 0.f;
#pragma unroll
      for (int b = 0; b < 4; ++b) aW2[a][b] = splat(0.f);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) aW1[a][b] = splat(0.f);
#pragma unroll
    for (int s = 0; s < kO; ++s) ab3[s] = 0.f;
  }

  const int64_t ntiles = (A.n + kTile - 1) / kTile;
  for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < ntiles;
       t += (int64_t)gridDim.x * kWaves) {
    // ---- per-row features (lane = row)
    const int64_t i = t * kTile + lane;
    const bool valid = i < A.n;
    int pos[kO];
    double di = 1.0;
#pragma unroll
    for (int s = 0; s < kO; ++s) pos[s] = -1;
    float x[kF];
#pragma unroll
    for (int f = 0; f < kF; ++f) x[f] = 0.f;
    if (valid) {
      row_slots(A.rp, A.ci, i, pos);
      di = A.diag[i];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int p = pos[k < 4 ? k : k + 1];
        if (p >= 0) {
          const int j = A.ci[p];
          const double dj = A.diag[j];
          x[k] = (float)(A.v[p] / di);
          x[8 + k] = (float)((dj - di) / (dj + di));
          x[16 + k] = (float)(A.rsum[j] / dj);
        }
      }
    }
#pragma unroll
    for (int f = 0; f < kFp; ++f) X[lane * kFp + f] = to_bf16(f < kF ? x[f] : 0.f);
    __builtin_amdgcn_wave_barrier();

    // ---- forward
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        floatx4 acc = splat(sw.b1[nt * 16 + r]);
        acc = mma(X + mt * 16 * kFp, kFp, 1, sw.w1 + nt * 16, kH, 1, acc);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          H1[(mt * 16 + 4 * g + e) * kH + nt * 16 + r] = to_bf16(fmaxf(acc[e], 0.f));
      }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        floatx4 acc = splat(sw.b2[nt * 16 + r]);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          acc = mma(H1 + mt * 16 * kH + kb * 32, kH, 1,
                    sw.w2 + kb * 32 * kH + nt * 16, kH, 1, acc);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          H2[(mt * 16 + 4 * g + e) * kH + nt * 16 + r] = to_bf16(fmaxf(acc[e], 0.f));
      }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = splat(sw.b3[r]);
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
        acc = mma(H2 + mt * 16 * kH + kb * 32, kH, 1, sw.w3 + kb * 32 * kOk,
                  kOk, 1, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) Y[(mt * 16 + 4 * g + e) * kOn + r] = acc[e];
    }
    __builtin_amdgcn_wave_barrier();
    float y[kO];
#pragma unroll
    for (int s = 0; s < kO; ++s) y[s] = Y[lane * kOn + s];
    double m[kO];
#pragma unroll
    for (int s = 0; s < kO; ++s)
      m[s] = pos[s] < 0 ? 0.0 : ((s == 4 ? 1.0 : 0.0) + (double)y[s]) / di;

    if (!TRAIN) {
      if (valid) {
#pragma unroll
        for (int s = 0; s < kO; ++s)
          if (pos[s] >= 0) A.mvals[pos[s]] = m[s];
      }
      __builtin_amdgcn_wave_barrier();
      continue;
    }

    // ---- (M A)_i by a merge of the slot rows; g_s = dL/dm_s (or both
    // from the symmetric objective's passes)
    double gs[kO];
#pragma unroll
    for (int s = 0; s < kO; ++s) gs[s] = 0.0;
    double lrow = 0.0;
    if (A.gm) {
      if (valid) {
#pragma unroll
        for (int s = 0; s < kO; ++s) gs[s] = pos[s] >= 0 ? A.gm[pos[s]] : 0.0;
        lrow = A.rowloss[i];
      }
    } else {
      lrow = row_residual(A.rp, A.ci, A.v, i, pos, m, gs);
    }
    if (valid) lsum += lrow;
    float dy[kO];
#pragma unroll
    for (int s = 0; s < kO; ++s) {
      dy[s] = (valid && pos[s] >= 0) ? (float)(A.scale * gs[s] / di) : 0.f;
      ab3[s] += dy[s];
    }
    // lane's Y row is read; its dY row occupies the same bytes
#pragma unroll
    for (int s = 0; s < kOk; ++s) dY[lane * kOk + s] = to_bf16(s < kO ? dy[s] : 0.f);
    __builtin_amdgcn_wave_barrier();

    // ---- backward
    // dW3[64 hid][16 out] += H2^T dY
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
        aW3[mt] = mma(H2 + mt * 16 + kb * 32 * kH, 1, kH, dY + kb * 32 * kOk,
                      kOk, 1, aW3[mt]);
    // dH2 = dY W3^T * [H2 > 0] -> D
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        floatx4 acc = mma(dY + mt * 16 * kOk, kOk, 1, sw.w3 + nt * 16 * kOk, 1,
                          kOk, splat(0.f));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int idx = (mt * 16 + 4 * g + e) * kH + nt * 16 + r;
          const float vv = from_bf16(H2[idx]) > 0.f ? acc[e] : 0.f;
          ab2[nt] += vv;
          D[idx] = to_bf16(vv);
        }
      }
    __builtin_amdgcn_wave_barrier();
    // dW2[64 in][64 out] += H1^T dH2
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          aW2[mt][nt] = mma(H1 + mt * 16 + kb * 32 * kH, 1, kH,
                            D + kb * 32 * kH + nt * 16, kH, 1, aW2[mt][nt]);
    // dH1 = dH2 W2^T * [H1 > 0] -> H2's region (H2 is no longer read)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        floatx4 acc = splat(0.f);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          acc = mma(D + mt * 16 * kH + kb * 32, kH, 1, sw.w2 + nt * 16 * kH + kb * 32,
                    1, kH, acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int idx = (mt * 16 + 4 * g + e) * kH + nt * 16 + r;
          const float vv = from_bf16(H1[idx]) > 0.f ? acc[e] : 0.f;
          ab1[nt] += vv;
          H2[idx] = to_bf16(vv);
        }
      }
    __builtin_amdgcn_wave_barrier();
    // dW1[32 feat][64 hid] += X^T dH1
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          aW1[mt][nt] = mma(X + mt * 16 + kb * 32 * kFp, 1, kFp,
                            H2 + kb * 32 * kH + nt * 16, kH, 1, aW1[mt][nt]);
    __builtin_amdgcn_wave_barrier();
  }
  if (!TRAIN) return;

  // ---- workgroup sum of the wavefronts' gradients, in a fixed order:
  // wavefront by wavefront, and for values several lanes of one wavefront
  // hold, lane group by lane group (every step behind a workgroup barrier)
  __syncthreads();
  float* red = (float*)pool;                        // kGradLen floats
  double* lt = (double*)(pool + 2 * kGradLen + 8);  // kWaves * 64 doubles
  for (int w = 0; w < kWaves; ++w) {
    if (wave == w) {
      const bool first = w == 0;
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int idx = gW1 + (mt * 16 + 4 * g + e) * kH + nt * 16 + r;
            red[idx] = (first ? 0.f : red[idx]) + aW1[mt][nt][e];
          }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int idx = gW2 + (mt * 16 + 4 * g + e) * kH + nt * 16 + r;
            red[idx] = (first ? 0.f : red[idx]) + aW2[mt][nt][e];
          }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int idx = gW3 + (mt * 16 + 4 * g + e) * kOn + r;
          red[idx] = (first ? 0.f : red[idx]) + aW3[mt][e];
        }
#pragma unroll
      for (int s = 0; s < kO; ++s) b3tmp[lane * kO + s] = ab3[s];
      lt[threadIdx.x] = lsum;
    }
    for (int gg = 0; gg < 4; ++gg) {
      __syncthreads();
      if (wave == w && g == gg) {
        const bool z = w == 0 && gg == 0;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
          const int i1 = gB1 + nt * 16 + r, i2 = gB2 + nt * 16 + r;
          red[i1] = (z ? 0.f : red[i1]) + ab1[nt];
          red[i2] = (z ? 0.f : red[i2]) + ab2[nt];
        }
      }
    }
    __syncthreads();
    if (threadIdx.x < kOn) {
      float acc = w == 0 ? 0.f : red[gB3 + threadIdx.x];
      if (threadIdx.x < kO)
        for (int l = 0; l < kTile; ++l) acc += b3tmp[l * kO + threadIdx.x];
      red[gB3 + threadIdx.x] = acc;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double l = 0.0;
    for (int k = 0; k < kWaves * 64; ++k) l += lt[k];
    A.loss_part[blockIdx.x] =
        (A.accumulate ? A.loss_part[blockIdx.x] : 0.0) + A.scale * l;
  }
  float* dst = A.partials + (size_t)blockIdx.x * kGradLen;
  for (int k = threadIdx.x; k < kGradLen; k += blockDim.x)
    dst[k] = (A.accumulate ? dst[k] : 0.f) + red[k];
}

// ---- the symmetric objective (1/n) ||I - S A||_F^2, S = (M + M^T) / 2:
// with mirror(p) the position of entry (j, i) for entry p = (i, j),
// S_p = (m_p + m_mirror(p)) / 2, each row's residual and g_p = dL/dS_p come
// from row_residual over S's values, and dL/dm_p = (g_p + g_mirror(p)) / 2
// (an entry of M enters S twice, at p and at its mirror).

// mirror positions, one thread per row; a missing mirror (a pattern that
// is not structurally symmetric) sets *err
__global__ void spai_mirror_kernel(int64_t n, const int32_t* __restrict__ rp,
                                   const int32_t* __restrict__ ci,
                                   int32_t* __restrict__ mirror, int* err) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int p = rp[i]; p < rp[i + 1]; ++p) {
    const int j = ci[p];
    int lo = rp[j], hi = rp[j + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ci[mid] < (int32_t)i) lo = mid + 1;
      else hi = mid;
    }
    const bool ok = lo < rp[j + 1] && ci[lo] == (int32_t)i;
    if (!ok) atomicOr(err, 1);
    mirror[p] = ok ? lo : p;
  }
}

// out_p = (in_p + in_mirror(p)) / 2 over the entries
__global__ void spai_mirror_mean_kernel(int64_t nnz, const double* __restrict__ in,
                                        const int32_t* __restrict__ mirror,
                                        double* __restrict__ out) {
#pragma clang fp contract(off)
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz;
       p += (int64_t)gridDim.x * blockDim.x)
    out[p] = 0.5 * (in[p] + in[mirror[p]]);
}

// each row's residual under S's values: the row's loss and g_p = dL/dS_p
__global__ void spai_row_loss_kernel(int64_t n, const int32_t* __restrict__ rp,
                                     const int32_t* __restrict__ ci,
                                     const double* __restrict__ v,
                                     const double* __restrict__ svals,
                                     double* __restrict__ g,
                                     double* __restrict__ rowloss) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int pos[kO];
  row_slots(rp, ci, i, pos);
  double m[kO], gs[kO];
#pragma unroll
  for (int s = 0; s < kO; ++s) {
    m[s] = pos[s] >= 0 ? svals[pos[s]] : 0.0;
    gs[s] = 0.0;
  }
  rowloss[i] = row_residual(rp, ci, v, i, pos, m, gs);
#pragma unroll
  for (int s = 0; s < kO; ++s)
    if (pos[s] >= 0) g[pos[s]] = gs[s];
}

// canonical parameter k -> (gradient partial index, packed slot)
__device__ __forceinline__ int partial_index(int k) {
  if (k < kB1) return gW1 + k;  // [f][o], same row length 64
  if (k < kW2) return gB1 + (k - kB1);
  if (k < kB2) return gW2 + (k - kW2);
  if (k < kW3) return gB2 + (k - kB2);
  if (k < kB3) return gW3 + ((k - kW3) / kO) * kOn + (k - kW3) % kO;
  return gB3 + (k - kB3);
}

__device__ __forceinline__ void store_packed(Packed* pk, int k, float w) {
  if (k < kB1) pk->w1[k] = to_bf16(w);
  else if (k < kW2) pk->b1[k - kB1] = w;
  else if (k < kB2) pk->w2[k - kW2] = to_bf16(w);
  else if (k < kW3) pk->b2[k - kB2] = w;
  else if (k < kB3) pk->w3[((k - kW3) / kO) * kOk + (k - kW3) % kO] = to_bf16(w);
  else pk->b3[k - kB3] = w;
}

struct AdamStep {
  float lr, beta1, beta2, eps, bc1, bc2;  // bc = 1 - beta^t
};

// sum the workgroup partials (in workgroup order); either write the
// gradient (grad != null) or take an Adam step on the master weights
__global__ void spai_update_kernel(int nparts, const float* __restrict__ partials,
                                   const double* __restrict__ loss_part,
                                   float* __restrict__ master, float* __restrict__ m1,
                                   float* __restrict__ m2, Packed* __restrict__ pk,
                                   double* __restrict__ loss_out,
                                   float* __restrict__ grad, AdamStep st) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    double l = 0.0;
    for (int b = 0; b < nparts; ++b) l += loss_part[b];
    *loss_out = l;
  }
  if (k >= kNP) return;
  const int pi = partial_index(k);
  float gsum = 0.f;
  for (int b = 0; b < nparts; ++b) gsum += partials[(size_t)b * kGradLen + pi];
  if (grad) {
    grad[k] = gsum;
    return;
  }
  const float a = st.beta1 * m1[k] + (1.f - st.beta1) * gsum;
  const float b = st.beta2 * m2[k] + (1.f - st.beta2) * (gsum * gsum);
  m1[k] = a;
  m2[k] = b;
  const float w = master[k] - st.lr * (a / st.bc1) / (sqrtf(b / st.bc2) + st.eps);
  master[k] = w;
  store_packed(pk, k, w);
}

__global__ void spai_pack_kernel(const float* __restrict__ master,
                                 Packed* __restrict__ pk) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kNP) store_packed(pk, k, master[k]);
}

}  // namespace
}  // namespace spk

using namespace spk;

struct spk_spai_model_s {
  DevBuf<float> master, m1, m2;
  DevBuf<Packed> pk;
  DevBuf<float> partials;
  DevBuf<double> loss_part, loss_hist;
  DevBuf<float> grad;
  DevBuf<double> diag, rsum;
  DevBuf<int> err;
  int64_t step = 0;
  int nparts = 0;
  // training objective: SPK_SPAI_LOSS_INVERSE (||I - M A||) or
  // SPK_SPAI_LOSS_SYMMETRIC (||I - S A||, S = (M + M^T) / 2), with the
  // latter's per-matrix scratch on A's pattern
  int objective = SPK_SPAI_LOSS_INVERSE;
  DevBuf<int32_t> mirror;
  DevBuf<double> mv, sv, gsv, gmv, rowl;

  void reset_optimizer() {
    hipStream_t s = stream();
    SPK_HIP_TRY(hipMemsetAsync(m1.p, 0, kNP * 4, s));
    SPK_HIP_TRY(hipMemsetAsync(m2.p, 0, kNP * 4, s));
    step = 0;
  }
  void upload(const float* host) {
    hipStream_t s = stream();
    SPK_HIP_TRY(hipMemcpyAsync(master.p, host, kNP * 4, hipMemcpyHostToDevice, s));
    SPK_HIP_TRY(hipMemsetAsync(pk.p, 0, sizeof(Packed), s));
    hipLaunchKernelGGL(spai_pack_kernel, dim3((kNP + 255) / 256), dim3(256), 0, s,
                       (const float*)master.p, pk.p);
    SPK_HIP_TRY(hipGetLastError());
    reset_optimizer();
    SPK_HIP_TRY(hipStreamSynchronize(s));
  }
  // diagonal / row sums of A into the model's scratch; throws on a matrix
  // the slot rules reject
  void prepare(spk_csr_s& A, bool check) {
    SPK_REQUIRE(A.nrows == A.ncols, SPK_INVALID_ARGUMENT,
                "learned SPAI needs a square matrix");
    const size_t n = (size_t)A.nrows;
    if (diag.n < n) {
      diag.alloc(n);
      rsum.alloc(n);
    }
    hipStream_t s = stream();
    if (check) SPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(spai_prepare_kernel, dim3(ceil_div(A.nrows, 256)), dim3(256),
                       0, s, A.nrows, (const int32_t*)A.rp.p,
                       (const int32_t*)A.ci.p, (const double*)A.v64.p, diag.p,
                       rsum.p, err.p);
    SPK_HIP_TRY(hipGetLastError());
    if (check) {
      int h = 0;
      SPK_HIP_TRY(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      SPK_REQUIRE(!h, SPK_INVALID_ARGUMENT,
                  "learned SPAI needs a positive diagonal and at most 4 "
                  "entries on each side of it in every row");
    }
  }
  TileArgs args(spk_csr_s& A) {
    TileArgs t{};
    t.n = A.nrows;
    t.rp = A.rp.p;
    t.ci = A.ci.p;
    t.v = A.v64.p;
    t.diag = diag.p;
    t.rsum = rsum.p;
    t.pk = pk.p;
    return t;
  }
  // grid of a predicting launch over A's row tiles
  int predict_grid(const spk_csr_s& A) const {
    const int64_t ntiles = (A.nrows + kTile - 1) / kTile;
    const int64_t want = (ntiles + kWaves - 1) / kWaves;
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, nparts));
  }
  // M's values on A's pattern (every entry is a slot: prepare checked)
  void predict_into(spk_csr_s& A, double* mvals) {
    TileArgs t = args(A);
    t.mvals = mvals;
    hipLaunchKernelGGL(spai_tile_kernel<false>, dim3(predict_grid(A)),
                       dim3(kWaves * 64), 0, stream(), t);
    SPK_HIP_TRY(hipGetLastError());
  }
  // The symmetric objective's passes before the backward one: M, the
  // mirrors, S, each row's loss and dL/dS, then dL/dm (into gmv / rowl).
  // check: report a pattern that is not structurally symmetric.
  void symmetric_scratch(const spk_csr_s& A) {
    const size_t nz = (size_t)(A.nnz > 0 ? A.nnz : 1), n = (size_t)A.nrows;
    if (mirror.n < nz) {
      mirror.alloc(nz);
      mv.alloc(nz);
      sv.alloc(nz);
      gsv.alloc(nz);
      gmv.alloc(nz);
    }
    if (rowl.n < n) rowl.alloc(n);
  }
  // the mirror positions of A's entries; check: SPK_INVALID_ARGUMENT for a
  // pattern that is not structurally symmetric
  void mirrors(const spk_csr_s& A, bool check) {
    symmetric_scratch(A);
    hipStream_t s = stream();
    if (check) SPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(spai_mirror_kernel, dim3((int)ceil_div(A.nrows, 256)),
                       dim3(256), 0, s, A.nrows, (const int32_t*)A.rp.p,
                       (const int32_t*)A.ci.p, mirror.p, err.p);
    SPK_HIP_TRY(hipGetLastError());
    if (check) {
      int h = 0;
      SPK_HIP_TRY(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      SPK_REQUIRE(!h, SPK_INVALID_ARGUMENT,
                  "the symmetric objective needs a structurally symmetric pattern");
    }
  }
  void symmetric_grad(spk_csr_s& A) {
    hipStream_t s = stream();
    predict_into(A, mv.p);
    mirrors(A, false);
    const int gr = (int)ceil_div(A.nrows, 256);
    const int ge = (int)std::min<int64_t>(ceil_div(A.nnz > 0 ? A.nnz : 1, 256), 65536);
    hipLaunchKernelGGL(spai_mirror_mean_kernel, dim3(ge), dim3(256), 0, s, A.nnz,
                       (const double*)mv.p, (const int32_t*)mirror.p, sv.p);
    hipLaunchKernelGGL(spai_row_loss_kernel, dim3(gr), dim3(256), 0, s, A.nrows,
                       (const int32_t*)A.rp.p, (const int32_t*)A.ci.p,
                       (const double*)A.v64.p, (const double*)sv.p, gsv.p, rowl.p);
    hipLaunchKernelGGL(spai_mirror_mean_kernel, dim3(ge), dim3(256), 0, s, A.nnz,
                       (const double*)gsv.p, (const int32_t*)mirror.p, gmv.p);
    SPK_HIP_TRY(hipGetLastError());
  }
  void train_tiles(spk_csr_s& A, double scale, bool accumulate) {
    TileArgs t = args(A);
    if (objective == SPK_SPAI_LOSS_SYMMETRIC) {
      symmetric_grad(A);
      t.gm = gmv.p;
      t.rowloss = rowl.p;
    }
    t.partials = partials.p;
    t.loss_part = loss_part.p;
    t.scale = scale;
    t.accumulate = accumulate ? 1 : 0;
    hipLaunchKernelGGL(spai_tile_kernel<true>, dim3(nparts), dim3(kWaves * 64), 0,
                       stream(), t);
    SPK_HIP_TRY(hipGetLastError());
  }
};

namespace {

uint64_t host_mix(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double host_u01(uint64_t z) {
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// a store with A's pattern and plan, values uninitialised
spk_csr_s* clone_pattern(const spk_csr_s& A) {
  auto* M = new spk_csr_s();
  try {
    M->allocate(A.nrows, A.ncols, A.nnz);
    hipStream_t s = stream();
    SPK_HIP_TRY(hipMemcpyAsync(M->rp.p, A.rp.p, (A.nrows + 1) * 4,
                               hipMemcpyDeviceToDevice, s));
    if (A.nnz)
      SPK_HIP_TRY(hipMemcpyAsync(M->ci.p, A.ci.p, A.nnz * 4,
                                 hipMemcpyDeviceToDevice, s));
    M->h_rp = A.h_rp;
    M->plan.alloc((size_t)A.nblocks + 1);
    SPK_HIP_TRY(hipMemcpyAsync(M->plan.p, A.plan.p, (A.nblocks + 1) * sizeof(int2),
                               hipMemcpyDeviceToDevice, s));
    M->nblocks = A.nblocks;
    M->nlong = A.nlong;
  } catch (...) {
    delete M;
    throw;
  }
  return M;
}

}  // namespace

extern "C" {

spk_status spk_spai_model_create(uint64_t seed, spk_spai_model* out) {
  return guarded([&] {
    SPK_REQUIRE(out, SPK_INVALID_ARGUMENT, "null argument");
    auto* m = new spk_spai_model_s();
    try {
      m->master.alloc(kNP);
      m->m1.alloc(kNP);
      m->m2.alloc(kNP);
      m->grad.alloc(kNP);
      m->pk.alloc(1);
      m->nparts = ctx().num_cus > 0 ? ctx().num_cus : 256;
      m->partials.alloc((size_t)m->nparts * kGradLen);
      m->loss_part.alloc((size_t)m->nparts);
      m->loss_hist.alloc(1);
      m->err.alloc(1);
      // He-uniform hidden layers, zero output layer (M starts as Jacobi)
      std::vector<float> p(kNP, 0.f);
      const uint64_t s = seed ^ 0x4D4C50ULL;
      for (int k = kW1; k < kB1; ++k)
        p[k] = (float)((2.0 * host_u01(host_mix(s, (uint64_t)k)) - 1.0) *
                       std::sqrt(6.0 / kF));
      for (int k = kW2; k < kB2; ++k)
        p[k] = (float)((2.0 * host_u01(host_mix(s, (uint64_t)k)) - 1.0) *
                       std::sqrt(6.0 / kH));
      m->upload(p.data());
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

spk_status spk_spai_model_destroy(spk_spai_model m) {
  return guarded([&] {
    if (!m) return;
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    delete m;
  });
}

spk_status spk_spai_model_get_params(spk_spai_model m, float* params) {
  return guarded([&] {
    SPK_REQUIRE(m && params, SPK_INVALID_ARGUMENT, "null argument");
    SPK_HIP_TRY(hipMemcpyAsync(params, m->master.p, kNP * 4,
                               hipMemcpyDeviceToHost, stream()));
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
  });
}

spk_status spk_spai_model_set_params(spk_spai_model m, const float* params) {
  return guarded([&] {
    SPK_REQUIRE(m && params, SPK_INVALID_ARGUMENT, "null argument");
    m->upload(params);
  });
}

// Checkpoint file (little-endian): "SPKSPAI" + NUL, uint32 version (1),
// uint32 parameter count, int64 Adam step, then the FP32 master weights,
// Adam's first and second moments (SPK_SPAI_NPARAM floats each, canonical
// order), then a 64-bit FNV-1a hash of everything before it.
namespace {
constexpr char kCkptMagic[8] = {'S', 'P', 'K', 'S', 'P', 'A', 'I', '\0'};
constexpr uint32_t kCkptVersion = 1;
constexpr size_t kCkptHead = 8 + 4 + 4 + 8;
constexpr size_t kCkptBytes = kCkptHead + 3 * (size_t)kNP * 4 + 8;

uint64_t fnv1a(const unsigned char* p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
  return h;
}
}  // namespace

spk_status spk_spai_model_save(spk_spai_model m, const char* path) {
  return guarded([&] {
    SPK_REQUIRE(m && path, SPK_INVALID_ARGUMENT, "null argument");
    std::vector<unsigned char> buf(kCkptBytes);
    unsigned char* q = buf.data();
    std::memcpy(q, kCkptMagic, 8);
    std::memcpy(q + 8, &kCkptVersion, 4);
    const uint32_t np = (uint32_t)kNP;
    std::memcpy(q + 12, &np, 4);
    std::memcpy(q + 16, &m->step, 8);
    hipStream_t s = stream();
    const DevBuf<float>* parts[3] = {&m->master, &m->m1, &m->m2};
    for (int k = 0; k < 3; ++k)
      SPK_HIP_TRY(hipMemcpyAsync(q + kCkptHead + (size_t)k * kNP * 4, parts[k]->p,
                                 (size_t)kNP * 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    const uint64_t h = fnv1a(q, kCkptBytes - 8);
    std::memcpy(q + kCkptBytes - 8, &h, 8);
    FILE* f = std::fopen(path, "wb");
    SPK_REQUIRE(f, SPK_RUNTIME_ERROR, std::string("cannot open ") + path);
    const size_t w = std::fwrite(q, 1, kCkptBytes, f);
    const int c = std::fclose(f);
    SPK_REQUIRE(w == kCkptBytes && c == 0, SPK_RUNTIME_ERROR,
                std::string("cannot write ") + path);
  });
}

spk_status spk_spai_model_load(spk_spai_model m, const char* path) {
  return guarded([&] {
    SPK_REQUIRE(m && path, SPK_INVALID_ARGUMENT, "null argument");
    FILE* f = std::fopen(path, "rb");
    SPK_REQUIRE(f, SPK_RUNTIME_ERROR, std::string("cannot open ") + path);
    std::vector<unsigned char> buf(kCkptBytes + 1);
    const size_t r = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    const unsigned char* q = buf.data();
    SPK_REQUIRE(r >= 16 && std::memcmp(q, kCkptMagic, 8) == 0, SPK_RUNTIME_ERROR,
                std::string(path) + " is not a learned-SPAI checkpoint");
    uint32_t ver = 0, np = 0;
    std::memcpy(&ver, q + 8, 4);
    std::memcpy(&np, q + 12, 4);
    SPK_REQUIRE(ver == kCkptVersion && np == (uint32_t)kNP, SPK_RUNTIME_ERROR,
                std::string(path) + ": unsupported checkpoint version or size");
    SPK_REQUIRE(r == kCkptBytes, SPK_RUNTIME_ERROR,
                std::string(path) + ": checkpoint has the wrong length");
    uint64_t h = 0;
    std::memcpy(&h, q + kCkptBytes - 8, 8);
    SPK_REQUIRE(h == fnv1a(q, kCkptBytes - 8), SPK_RUNTIME_ERROR,
                std::string(path) + ": checkpoint checksum mismatch");
    int64_t step = 0;
    std::memcpy(&step, q + 16, 8);
    SPK_REQUIRE(step >= 0, SPK_RUNTIME_ERROR, std::string(path) + ": bad step");
    // weights (and their packed copy) first, then the optimiser state that
    // upload() resets
    m->upload(reinterpret_cast<const float*>(q + kCkptHead));
    hipStream_t s = stream();
    SPK_HIP_TRY(hipMemcpyAsync(m->m1.p, q + kCkptHead + (size_t)kNP * 4,
                               (size_t)kNP * 4, hipMemcpyHostToDevice, s));
    SPK_HIP_TRY(hipMemcpyAsync(m->m2.p, q + kCkptHead + 2 * (size_t)kNP * 4,
                               (size_t)kNP * 4, hipMemcpyHostToDevice, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    m->step = step;
  });
}

spk_status spk_spai_model_set_objective(spk_spai_model m, int objective) {
  return guarded([&] {
    SPK_REQUIRE(m, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(objective == SPK_SPAI_LOSS_INVERSE ||
                    objective == SPK_SPAI_LOSS_SYMMETRIC,
                SPK_INVALID_ARGUMENT, "unknown objective");
    m->objective = objective;
  });
}

spk_status spk_spai_model_get_objective(spk_spai_model m, int* objective) {
  return guarded([&] {
    SPK_REQUIRE(m && objective, SPK_INVALID_ARGUMENT, "null argument");
    *objective = m->objective;
  });
}

spk_status spk_spai_features(spk_csr A, float* d_features) {
  return guarded([&] {
    SPK_REQUIRE(A && d_features, SPK_INVALID_ARGUMENT, "null argument");
    // computed by the same per-row code as the tile kernel, through a model
    // whose forward pass is skipped: a dedicated small kernel
    spk_spai_model_s tmp;
    tmp.err.alloc(1);
    tmp.prepare(*A, true);
    const int32_t* rp = A->rp.p;
    const int32_t* ci = A->ci.p;
    const double* v = A->v64.p;
    const double* diag = tmp.diag.p;
    const double* rsum = tmp.rsum.p;
    float* X = d_features;
    launch_map(A->nrows, [=] __device__(int64_t i) {
      const int p0 = rp[i], p1 = rp[i + 1];
      int d = p0;
      while (ci[d] != (int32_t)i) ++d;
      int pos[kO];
      pos[4] = d;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        pos[3 - k] = (d - 1 - k >= p0) ? d - 1 - k : -1;
        pos[5 + k] = (d + 1 + k < p1) ? d + 1 + k : -1;
      }
      const double di = diag[i];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int p = pos[k < 4 ? k : k + 1];
        float a = 0.f, b = 0.f, c = 0.f;
        if (p >= 0) {
          const int j = ci[p];
          const double dj = diag[j];
          a = (float)(v[p] / di);
          b = (float)((dj - di) / (dj + di));
          c = (float)(rsum[j] / dj);
        }
        X[i * kF + k] = a;
        X[i * kF + 8 + k] = b;
        X[i * kF + 16 + k] = c;
      }
    });
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
  });
}

spk_status spk_spai_loss_grad(spk_spai_model m, spk_csr A, double* loss,
                              float* grad) {
  return guarded([&] {
    SPK_REQUIRE(m && A && loss, SPK_INVALID_ARGUMENT, "null argument");
    m->prepare(*A, true);
    if (m->objective == SPK_SPAI_LOSS_SYMMETRIC) m->mirrors(*A, true);
    m->train_tiles(*A, 1.0 / (double)A->nrows, false);
    hipStream_t s = stream();
    hipLaunchKernelGGL(spai_update_kernel, dim3((kNP + 255) / 256), dim3(256), 0,
                       s, m->nparts, (const float*)m->partials.p,
                       (const double*)m->loss_part.p, m->master.p, m->m1.p,
                       m->m2.p, m->pk.p, m->loss_hist.p, m->grad.p, AdamStep{});
    SPK_HIP_TRY(hipGetLastError());
    SPK_HIP_TRY(hipMemcpyAsync(loss, m->loss_hist.p, 8, hipMemcpyDeviceToHost, s));
    if (grad)
      SPK_HIP_TRY(hipMemcpyAsync(grad, m->grad.p, kNP * 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
  });
}

spk_status spk_spai_train(spk_spai_model m, const spk_csr* mats, int64_t count,
                          const spk_spai_train_config* cfg,
                          double* loss_history) {
  return guarded([&] {
    TraceRange trace_("spk_spai_train");
    SPK_REQUIRE(m && mats && cfg && count > 0, SPK_INVALID_ARGUMENT,
                "null argument");
    SPK_REQUIRE(cfg->epochs >= 0 && cfg->batch_matrices > 0 &&
                    cfg->learning_rate > 0.0,
                SPK_INVALID_ARGUMENT, "bad training configuration");
    for (int64_t k = 0; k < count; ++k) {
      SPK_REQUIRE(mats[k], SPK_INVALID_ARGUMENT, "null matrix");
      m->prepare(*mats[k], true);
      if (m->objective == SPK_SPAI_LOSS_SYMMETRIC) m->mirrors(*mats[k], true);
    }
    const int64_t per_epoch = (count + cfg->batch_matrices - 1) / cfg->batch_matrices;
    const int64_t total = cfg->epochs * per_epoch;
    if (m->loss_hist.n < (size_t)(total > 0 ? total : 1))
      m->loss_hist.alloc((size_t)(total > 0 ? total : 1));
    hipStream_t s = stream();
    std::vector<int64_t> order((size_t)count);
    const float b1 = 0.9f, b2 = 0.999f;
    int64_t h = 0;
    for (int64_t e = 0; e < cfg->epochs; ++e) {
      for (int64_t k = 0; k < count; ++k) order[(size_t)k] = k;
      for (int64_t k = count - 1; k > 0; --k) {  // Fisher-Yates
        uint64_t z = host_mix(cfg->shuffle_seed, (uint64_t)(e * count + k));
        std::swap(order[(size_t)k], order[(size_t)(z % (uint64_t)(k + 1))]);
      }
      for (int64_t b0 = 0; b0 < count; b0 += cfg->batch_matrices) {
        const int64_t b1e = std::min(count, b0 + cfg->batch_matrices);
        int64_t rows = 0;
        for (int64_t k = b0; k < b1e; ++k) rows += mats[order[(size_t)k]]->nrows;
        for (int64_t k = b0; k < b1e; ++k) {
          spk_csr_s& A = *mats[order[(size_t)k]];
          m->prepare(A, false);
          m->train_tiles(A, 1.0 / (double)rows, k > b0);
        }
        m->step += 1;
        AdamStep st;
        st.lr = (float)cfg->learning_rate;
        st.beta1 = b1;
        st.beta2 = b2;
        st.eps = 1e-8f;
        st.bc1 = (float)(1.0 - std::pow((double)b1, (double)m->step));
        st.bc2 = (float)(1.0 - std::pow((double)b2, (double)m->step));
        hipLaunchKernelGGL(spai_update_kernel, dim3((kNP + 255) / 256), dim3(256),
                           0, s, m->nparts, (const float*)m->partials.p,
                           (const double*)m->loss_part.p, m->master.p, m->m1.p,
                           m->m2.p, m->pk.p, m->loss_hist.p + h, (float*)nullptr,
                           st);
        SPK_HIP_TRY(hipGetLastError());
        ++h;
      }
    }
    if (loss_history && total > 0)
      SPK_HIP_TRY(hipMemcpyAsync(loss_history, m->loss_hist.p, total * 8,
                                 hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
  });
}

spk_status spk_spai_predict(spk_spai_model m, spk_csr A, spk_csr* M_out) {
  return guarded([&] {
    TraceRange trace_("spk_spai_predict");
    SPK_REQUIRE(m && A && M_out, SPK_INVALID_ARGUMENT, "null argument");
    m->prepare(*A, true);
    spk_csr_s* M = clone_pattern(*A);
    try {
      m->predict_into(*A, M->v64.p);
      SPK_HIP_TRY(hipStreamSynchronize(stream()));
    } catch (...) {
      delete M;
      throw;
    }
    *M_out = M;
  });
}

}  // extern "C"
