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
