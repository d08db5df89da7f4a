/*
 * CPU restatement of the taichi_elements MLS-MPM substep (oracle) --
 * TEST INFRASTRUCTURE ONLY; see mpm_oracle.h for who may use it.
 * Compiled with -ffp-contract=off so every product and sum rounds once.
 */
#include "mpm_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

#define REAL float
#define SUF(x) x##_f32
#include "mpm_oracle_impl.inc"
#undef REAL
#undef SUF

#define REAL double
#define SUF(x) x##_f64
#include "mpm_oracle_impl.inc"
#undef REAL
#undef SUF

/* build_pid's block of a particle, mpm_solver.py:357-361. */
void oracle_block_keys(int dim, int64_t n, const float* x, float inv_dx,
                       int block_size, const int32_t* blo, const int32_t* bn,
                       int32_t* keys) {
  const int32_t invalid = bn[0] * bn[1] * (dim == 3 ? bn[2] : 1);
  for (int64_t p = 0; p < n; p++) {
    int32_t key = 0, ok = 1;
    for (int a = dim - 1; a >= 0; a--) {
      volatile float t = x[p * dim + a] * inv_dx;
      volatile float u = t - 0.5f;
      int base = (int)floorf(u);
      /* floor division by the block size */
      int b = base >= 0 ? base / block_size : -((-base + block_size - 1) / block_size);
      int lb = b - blo[a];
      if (lb < 0 || lb >= bn[a]) ok = 0;
      key = key * bn[a] + lb;
    }
    keys[p] = ok ? key : invalid;
  }
}

/* Stable counting sort (the reference appends indices to per-block lists,
 * mpm:361; this fixes the order inside a block to increasing index). */
void oracle_sort_by_key(int64_t n, const int32_t* keys, int32_t n_keys,
                        int32_t* perm) {
  int64_t* start = (int64_t*)calloc((size_t)n_keys + 2, sizeof(int64_t));
  for (int64_t p = 0; p < n; p++) start[keys[p] + 1]++;
  for (int32_t k = 0; k <= n_keys; k++) start[k + 1] += start[k];
  for (int64_t p = 0; p < n; p++) perm[start[keys[p]]++] = (int32_t)p;
  free(start);
}
