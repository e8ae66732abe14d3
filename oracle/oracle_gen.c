/*
 * oracle_gen.c -- CPU restatement of the generators that have no reference
 * counterpart (their definition is this build's; the device generators in
 * sparkit_amd/csrc/spk_gen.hip implement the same definition independently
 * and the tests compare the two bit for bit).  TEST INFRASTRUCTURE ONLY
 * (see sparkit_oracle.h).
 *
 * Power-law SPD matrix (n rows, seed):
 *   streams   sL = seed ^ 0x4C454E, sC = seed ^ 0x434F4C, sV = seed ^ 0x56414C
 *   u01(z)    = (z >> 11) * 2^-53
 *   lengths   L_i = min{k in [3,4096] : cdf[k] >= u01(mix(sL, i))},
 *             cdf[k] = S_k / S_4096, S_k = sum_{j=3..k} pow(j, -1.5)
 *             (a Zipf(1.5) law truncated to [3, 4096]; sequential sums)
 *   columns   draw t = 0..L_i-2: c = mix(sC, i*4096 + t) mod (n-1),
 *             c += (c >= i); duplicates collapse
 *   one-sided a(i,j) = -u01(mix(sV, i*n + j))           (uniform [-1, 0))
 *   pattern   union of the draws, their transposes and the diagonal
 *   values    A_ij = 0.5 * ([j drawn by i] a(i,j) + [i drawn by j] a(j,i))
 *             (the symmetric part (A + A^T)/2 on the union pattern)
 *   diagonal  A_ii = (sum over j != i, ascending j, of |A_ij|) + 1
 *             (sparkit's dominance rule, matgen.hpp:657-658)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sparkit_oracle.h"

#define PL_MIN 3
#define PL_MAX 4096

void orc_powerlaw_cdf(double* cdf /* PL_MAX - PL_MIN + 1 entries */) {
  double s = 0.0;
  for (int k = PL_MIN; k <= PL_MAX; ++k) {
    s += pow((double)k, -1.5);
    cdf[k - PL_MIN] = s;
  }
  double total = s;
  for (int k = PL_MIN; k <= PL_MAX; ++k) cdf[k - PL_MIN] /= total;
}

static double u01(uint64_t z) {
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int64_t orc_powerlaw_length(const double* cdf, uint64_t seed, int64_t i) {
  double u = u01(orc_splitmix64_at(seed ^ 0x4C454EULL, (uint64_t)i));
  int lo = 0, hi = PL_MAX - PL_MIN; /* first index with cdf >= u */
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (cdf[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  return lo + PL_MIN;
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

/* Pass NULL arrays to get nnz only (the pattern is built either way). */
int64_t orc_gen_powerlaw_spd(int64_t n, uint64_t seed, int64_t* row_ptr,
                             int64_t* col_ind, double* vals) {
  double* cdf = (double*)malloc(sizeof(double) * (PL_MAX - PL_MIN + 1));
  orc_powerlaw_cdf(cdf);
  int64_t draws = 0;
  for (int64_t i = 0; i < n; ++i) draws += orc_powerlaw_length(cdf, seed, i) - 1;
  /* keys: ((row * n + col) << 1) | dir, dir 0 = drawn by row, 1 = transposed */
  int64_t K = 2 * draws + n;
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)K);
  int64_t m = 0;
  const uint64_t un = (uint64_t)n;
  for (int64_t i = 0; i < n; ++i) {
    int64_t L = orc_powerlaw_length(cdf, seed, i);
    for (int64_t t = 0; t < L - 1; ++t) {
      uint64_t c = orc_splitmix64_at(seed ^ 0x434F4CULL,
                                     (uint64_t)i * PL_MAX + (uint64_t)t) %
                   (un - 1);
      if (c >= (uint64_t)i) c += 1;
      keys[m++] = (((uint64_t)i * un + c) << 1) | 0u;
      keys[m++] = ((c * un + (uint64_t)i) << 1) | 1u;
    }
    keys[m++] = (((uint64_t)i * un + (uint64_t)i) << 1) | 0u;
  }
  qsort(keys, (size_t)K, sizeof(uint64_t), cmp_u64);
  int64_t nnz = 0;
  if (row_ptr) row_ptr[0] = 0;
  int64_t cur_row = 0;
  for (int64_t k = 0; k < K;) {
    uint64_t pair = keys[k] >> 1;
    int has0 = (keys[k] & 1u) == 0;
    int64_t e = k;
    while (e < K && (keys[e] >> 1) == pair) ++e;
    int has1 = (keys[e - 1] & 1u) == 1;
    int64_t i = (int64_t)(pair / un), j = (int64_t)(pair % un);
    if (row_ptr) {
      while (cur_row < i) row_ptr[++cur_row] = nnz;
      col_ind[nnz] = j;
      if (i == j) {
        vals[nnz] = 0.0; /* diagonal: filled below */
      } else {
        double t = (has0 ? -u01(orc_splitmix64_at(seed ^ 0x56414CULL,
                                                  (uint64_t)i * un + (uint64_t)j))
                         : 0.0) +
                   (has1 ? -u01(orc_splitmix64_at(seed ^ 0x56414CULL,
                                                  (uint64_t)j * un + (uint64_t)i))
                         : 0.0);
        vals[nnz] = 0.5 * t;
      }
    }
    ++nnz;
    k = e;
  }
  if (row_ptr) {
    while (cur_row < n) row_ptr[++cur_row] = nnz;
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      int64_t dpos = -1;
      for (int64_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
        if (col_ind[p] == i) dpos = p;
        else s += fabs(vals[p]);
      }
      vals[dpos] = s + 1.0;
    }
  }
  free(keys);
  free(cdf);
  return nnz;
}

/*
 * 2D anisotropic variable-coefficient diffusion, bilinear (Q1) finite
 * elements on the unit square, homogeneous Dirichlet boundary eliminated
 * (N x N interior nodes, 9-point stencil, SPD).
 *
 *   streams   sP = seed ^ 0x414E49 (matrix parameters), sK = seed ^ 0x4B4150
 *   anisotropy  eps   = exp(ln(1000) * u01(mix(sP, 0)))     (log-uniform [1,1e3])
 *   angle       theta = pi * u01(mix(sP, 1))                (uniform [0, pi))
 *   tensor      K0 = R(theta) diag(1, eps) R(theta)^T:
 *               kxx = c c + eps s s, kyy = s s + eps c c, kxy = (1 - eps) c s
 *   cells     (cx, cy), 0 <= cx, cy <= N, id = cy (N+1) + cx,
 *             u1 = 1 - u01(mix(sK, 2 id)), u2 = u01(mix(sK, 2 id + 1)),
 *             z = sqrt(-2 log u1) cos(2 pi u2)  (Box-Muller),
 *             kappa = exp(sigma z)              (lognormal(0, sigma))
 *   element   corners q = dx + 2 dy of cell (cx, cy) are nodes
 *             (cx + dx, cy + dy); the exact element stiffness of the
 *             constant tensor kappa K0 on a square is kappa E with
 *             E[a][b] = (kxx X[a][b] + kyy Y[a][b] + kxy Z[a][b]) / 6,
 *             X, Y, Z the integer tables below (6 x the integrals of
 *             d_x phi_a d_x phi_b, d_y phi_a d_y phi_b and the symmetrised
 *             mixed term over the unit square)
 *   assembly  row p = (iy - 1) N + (ix - 1) for interior node (ix, iy);
 *             A(p, p') = sum over the cells holding both nodes, ascending
 *             cell id, of kappa_cell * E[q_p][q_p'] (summed from 0.0)
 * With sigma = 0 and eps = 1 this is the textbook Q1 Laplacian: 8/3 on the
 * diagonal, -1/3 for the eight neighbours.
 */
static const double kQ1X[4][4] = {
    {2, -2, 1, -1}, {-2, 2, -1, 1}, {1, -1, 2, -2}, {-1, 1, -2, 2}};
static const double kQ1Y[4][4] = {
    {2, 1, -2, -1}, {1, 2, -1, -2}, {-2, -1, 2, 1}, {-1, -2, 1, 2}};
static const double kQ1Z[4][4] = {
    {3, 0, 0, -3}, {0, -3, 3, 0}, {0, 3, -3, 0}, {-3, 0, 0, 3}};

void orc_aniso_q1_params(int64_t N, uint64_t seed, double sigma, int iso,
                         double* kappa, double* E) {
  double eps = 1.0, theta = 0.0;
  if (!iso) {
    eps = exp(6.907755278982137 * u01(orc_splitmix64_at(seed ^ 0x414E49ULL, 0)));
    theta = 3.141592653589793 * u01(orc_splitmix64_at(seed ^ 0x414E49ULL, 1));
  }
  const double c = cos(theta), s = sin(theta);
  const double kxx = c * c + eps * (s * s);
  const double kyy = s * s + eps * (c * c);
  const double kxy = (1.0 - eps) * (c * s);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      E[4 * a + b] =
          ((kxx * kQ1X[a][b] + kyy * kQ1Y[a][b]) + kxy * kQ1Z[a][b]) / 6.0;
  const int64_t nc = N + 1;
  for (int64_t id = 0; id < nc * nc; ++id) {
    if (sigma == 0.0) {
      kappa[id] = 1.0;
      continue;
    }
    double u1 = 1.0 - u01(orc_splitmix64_at(seed ^ 0x4B4150ULL, 2 * (uint64_t)id));
    double u2 = u01(orc_splitmix64_at(seed ^ 0x4B4150ULL, 2 * (uint64_t)id + 1));
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    kappa[id] = exp(sigma * z);
  }
}

int64_t orc_gen_aniso_q1_2d(int64_t N, uint64_t seed, double sigma, int iso,
                            int64_t* row_ptr, int64_t* col_ind, double* vals) {
  const int64_t nc = N + 1;
  double* kappa = (double*)malloc(sizeof(double) * (size_t)(nc * nc));
  double E[16];
  orc_aniso_q1_params(N, seed, sigma, iso, kappa, E);
  int64_t nnz = 0;
  row_ptr[0] = 0;
  for (int64_t iy = 1; iy <= N; ++iy)
    for (int64_t ix = 1; ix <= N; ++ix) {
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          int64_t jx = ix + ox, jy = iy + oy;
          if (jx < 1 || jx > N || jy < 1 || jy > N) continue;
          double t = 0.0;
          for (int b = 0; b <= 1; ++b)   /* cell row cy = iy - 1 + b */
            for (int a = 0; a <= 1; ++a) { /* cell column cx = ix - 1 + a */
              int64_t cx = ix - 1 + a, cy = iy - 1 + b;
              int64_t dxp = ix - cx, dyp = iy - cy, dxq = jx - cx, dyq = jy - cy;
              if (dxq < 0 || dxq > 1 || dyq < 0 || dyq > 1) continue;
              t += kappa[cy * nc + cx] *
                   E[4 * (dxp + 2 * dyp) + (dxq + 2 * dyq)];
            }
          col_ind[nnz] = (jy - 1) * N + (jx - 1);
          vals[nnz] = t;
          ++nnz;
        }
      row_ptr[(iy - 1) * N + ix] = nnz;
    }
  free(kappa);
  return nnz;
}

/*
 * Batched 2D variable-coefficient diffusion (cell-centred finite volumes,
 * 5-point, Dirichlet), system k of a batch (seed, side range [smin, smax]):
 *   side      s_k = smin + mix(seed ^ 0x534944, k) mod (smax - smin + 1)
 *   stream    t_k = mix(seed ^ 0x424154, k)
 *   cells     c = y s + x; u1 = 1 - u01(mix(t_k, 2c)),
 *             u2 = u01(mix(t_k, 2c + 1)), z = sqrt(-2 log u1) cos(2 pi u2),
 *             kappa_c = exp(z)                       (lognormal(0, 1))
 *   faces     interior: t = (2 kappa_a) kappa_b / (kappa_a + kappa_b)
 *             (harmonic mean; bitwise symmetric), boundary: t = 2 kappa_c
 *   row c     columns c - s, c - 1, c, c + 1, c + s (those inside), values
 *             -t of the face; diagonal ((t_S + t_W) + t_E) + t_N with the
 *             boundary faces included.
 */
int64_t orc_batched_side(uint64_t seed, int64_t k, int64_t smin, int64_t smax) {
  return smin + (int64_t)(orc_splitmix64_at(seed ^ 0x534944ULL, (uint64_t)k) %
                          (uint64_t)(smax - smin + 1));
}

void orc_batched_kappa(uint64_t seed, int64_t k, int64_t s, double* kappa) {
  const uint64_t t = orc_splitmix64_at(seed ^ 0x424154ULL, (uint64_t)k);
  for (int64_t c = 0; c < s * s; ++c) {
    double u1 = 1.0 - u01(orc_splitmix64_at(t, 2 * (uint64_t)c));
    double u2 = u01(orc_splitmix64_at(t, 2 * (uint64_t)c + 1));
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    kappa[c] = exp(z);
  }
}

static double face(double a, double b) { return (2.0 * a) * b / (a + b); }

int64_t orc_gen_batched_diffusion(uint64_t seed, int64_t k, int64_t smin,
                                  int64_t smax, int64_t* row_ptr,
                                  int64_t* col_ind, double* vals) {
  const int64_t s = orc_batched_side(seed, k, smin, smax);
  double* kap = (double*)malloc(sizeof(double) * (size_t)(s * s));
  orc_batched_kappa(seed, k, s, kap);
  int64_t nnz = 0;
  row_ptr[0] = 0;
  for (int64_t y = 0; y < s; ++y)
    for (int64_t x = 0; x < s; ++x) {
      const int64_t c = y * s + x;
      const double kc = kap[c];
      const double tS = y > 0 ? face(kc, kap[c - s]) : 2.0 * kc;
      const double tW = x > 0 ? face(kc, kap[c - 1]) : 2.0 * kc;
      const double tE = x < s - 1 ? face(kc, kap[c + 1]) : 2.0 * kc;
      const double tN = y < s - 1 ? face(kc, kap[c + s]) : 2.0 * kc;
      if (y > 0) { col_ind[nnz] = c - s; vals[nnz++] = -tS; }
      if (x > 0) { col_ind[nnz] = c - 1; vals[nnz++] = -tW; }
      col_ind[nnz] = c;
      vals[nnz++] = ((tS + tW) + tE) + tN;
      if (x < s - 1) { col_ind[nnz] = c + 1; vals[nnz++] = -tE; }
      if (y < s - 1) { col_ind[nnz] = c + s; vals[nnz++] = -tN; }
      row_ptr[c + 1] = nnz;
    }
  free(kap);
  return nnz;
}
