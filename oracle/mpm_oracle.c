/*
 * CPU restatement of the taichi_elements MLS-MPM substep (oracle) --
 * TEST INFRASTRUCTURE ONLY; see mpm_oracle.h for who may use it.
 * Compiled with -ffp-contract=off so every product and sum rounds once.
 */
#include "mpm_oracle.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

/* ------------------------------------------------------------------ */
/* Quantised particle fields (MPMSolver(quant=True), mpm_solver.py:106- */
/* 114): quant.fixed(21, max 2.0) for x, quant.float(exp=7, frac=19)    */
/* with a shared exponent for v, quant.fixed(16, max F_bound + 0.1) for */
/* F.  Taichi's quantised types are third-party (taichi==1.1.0, absent); */
/* their encodings are restated with round-to-nearest-even and          */
/* saturation, the engine's documented rounding.                        */
/* ------------------------------------------------------------------ */
static float quant_x1(float x) {
  float t = rintf(x * 524288.0f); /* x / 2^-19 */
  t = fminf(fmaxf(t, -1048576.0f), 1048575.0f);
  return (float)(int)t * (1.0f / 524288.0f); /* an integer code: -0 reads back as +0 */
}

static float quant_f1(float f) {
  const float inv = (float)(32768.0 / 4.1), step = (float)(4.1 / 32768.0);
  float t = rintf(f * inv);
  t = fminf(fmaxf(t, -32768.0f), 32767.0f);
  return (float)(int)t * step;
}

/* Shared exponent e with max |v_a| < 2^e; magnitudes |v_a| 2^(18-e)
 * rounded; e raised by one when the largest rounds up to 2^18; e kept in
 * [-62, 64] (smaller: the zero vector; larger or non-finite: saturated). */
static void quant_v1(int d, float* v) {
  float m = 0.0f;
  int finite = 1;
  for (int a = 0; a < d; a++) {
    const float u = fabsf(v[a]);
    if (!(u <= 3.4028234663852886e38f)) finite = 0;
    m = fmaxf(m, u);
  }
  float mag[3] = {0.0f, 0.0f, 0.0f};
  int e = 0, zero = 0;
  if (!finite) {
    e = 64;
    for (int a = 0; a < d; a++) {
      const float u = fabsf(v[a]);
      mag[a] = u <= 3.4028234663852886e38f ? rintf(ldexpf(u, 18 - 64)) : 262143.0f;
    }
  } else if (m > 0.0f) {
    frexpf(m, &e);
    if (e < -62) {
      zero = 1;
    } else {
      if (e > 64) e = 64;
      float big = 0.0f;
      for (int a = 0; a < d; a++) big = fmaxf(big, rintf(ldexpf(fabsf(v[a]), 18 - e)));
      if (big >= 262144.0f && e < 64) e++;
      for (int a = 0; a < d; a++) mag[a] = fminf(rintf(ldexpf(fabsf(v[a]), 18 - e)), 262143.0f);
    }
  } else {
    zero = 1;
  }
  for (int a = 0; a < d; a++) {
    if (zero) {
      v[a] = 0.0f;
      continue;
    }
    const float r = mag[a] * ldexpf(1.0f, e - 18);
    v[a] = (signbit(v[a]) && mag[a] != 0.0f) ? -r : r;
  }
}

void oracle_quant_x(int64_t n, float* x) {
  for (int64_t i = 0; i < n; i++) x[i] = quant_x1(x[i]);
}
void oracle_quant_f(int64_t n, float* f) {
  for (int64_t i = 0; i < n; i++) f[i] = quant_f1(f[i]);
}
void oracle_quant_v(int dim, int64_t n, float* v) {
  for (int64_t i = 0; i < n; i++) quant_v1(dim, v + i * dim);
}

#define REAL float
#define REAL_MIN_NORMAL FLT_MIN
#define SUF(x) x##_f32
#include "mpm_oracle_impl.inc"
#undef REAL
#undef REAL_MIN_NORMAL
#undef SUF

#define REAL double
#define REAL_MIN_NORMAL DBL_MIN
#define SUF(x) x##_f64
#include "mpm_oracle_impl.inc"
#undef REAL
#undef REAL_MIN_NORMAL
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
