// Particle storage: the float layout and the packed (quantised) layout of
// MPMSolver(quant=True) (mpm_solver.py:106-114, 216-262), behind one
// accessor type so that every kernel is written once for both.
//
// Both layouts are structure-of-arrays: component c of slot p is the 32-bit
// word at c * cap + p.
//
// Float layout (Layout<D>): x (D), v (D), C (D*D), F (D*D), Jp.
//
// Packed layout, word by word:
//   x   2 words: a 64-bit field of D signed 21-bit fixed-point values,
//       value = code * 2^-19 (range [-2, 2), the reference's
//       quant.fixed(21, max_value=2.0)), component a at bits [21a, 21a + 21).
//   v   2 words: a 64-bit field of one 7-bit shared exponent (bits 0-6) and
//       D 19-bit sign-magnitude mantissas (component a at bits
//       [7 + 19a, 26 + 19a), sign in the top bit): value = +-m * 2^(e - 18)
//       with e = E - 63 (E = 0: all zero) -- the reference's
//       quant.float(exp=7, frac=19) with a shared exponent.  The exponent is
//       the one of the largest component, so that component keeps 18 bits.
//   F   ceil(D*D / 2) words: signed 16-bit fixed-point values, value = code *
//       (4.1 / 32768) (quant.fixed(16, max_value=F_bound + 0.1)), two per
//       word, row-major (low half first).
//   C   D*D float words (absent in use_g2p2g mode, which keeps no C).
//   Jp  1 float word (absent with support_plasticity=False).
// Encoding rounds to nearest (ties to even) and saturates at the range
// ends; decoding is exact.  The material stays in the int components.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

// Float particle components, each a [cap] array inside one allocation:
// x (D), v (D), C (D*D), F (D*D), Jp (1).  Int components: material, slot
// insertion index, scene.
template <int D>
struct Layout {
  static constexpr int X = 0, V = D, C = 2 * D, F = 2 * D + D * D, JP = 2 * D + 2 * D * D;
  static constexpr int NF = 2 * D + 2 * D * D + 1;
  static constexpr int MAT = 0, ORIG = 1, SCENE = 2, NI = 3;
};

// Packed layout word offsets; G: use_g2p2g (no C), plastic: Jp present.
template <int D>
struct PackedLayout {
  static constexpr int X = 0, V = 2, F = 4, NFW = (D * D + 1) / 2;
  static constexpr int C = F + NFW;
  __host__ __device__ static constexpr int jp(bool g2p2g) { return C + (g2p2g ? 0 : D * D); }
  __host__ __device__ static constexpr int words(bool g2p2g, bool plastic) {
    return jp(g2p2g) + (plastic ? 1 : 0);
  }
};

constexpr float kFBound = 4.0f;                       // F_bound, mpm_solver.py:99
constexpr float kFQuantInv = (float)(32768.0 / 4.1);  // codes per unit of F
constexpr float kFQuantStep = (float)(4.1 / 32768.0);

__device__ __forceinline__ int quant_x_code(float x) {
  const float t = rintf(x * 524288.0f);  // x / 2^-19, exact scaling
  return (int)fminf(fmaxf(t, -1048576.0f), 1048575.0f);
}
__device__ __forceinline__ float quant_x_value(int code) { return (float)code * (1.0f / 524288.0f); }

__device__ __forceinline__ int quant_f_code(float f) {
#pragma clang fp contract(off)
  const float t = rintf(f * kFQuantInv);
  return (int)fminf(fmaxf(t, -32768.0f), 32767.0f);
}
__device__ __forceinline__ float quant_f_value(int code) {
#pragma clang fp contract(off)
  return (float)code * kFQuantStep;
}

// Shared-exponent encoding of D velocity components (see the header).
template <int D>
__device__ __forceinline__ uint64_t quant_v_encode(const float v[3]) {
  float m = 0.0f;
  bool finite = true;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const float u = fabsf(v[a]);
    finite &= u <= 3.4028234663852886e38f;  // false for inf and NaN
    m = fmaxf(m, u);
  }
  unsigned mag[3] = {0u, 0u, 0u};
  int E = 0;
  if (!finite) {
    E = 127;  // saturate: the largest representable magnitude
#pragma unroll
    for (int a = 0; a < D; a++) mag[a] = (fabsf(v[a]) <= 3.4028234663852886e38f)
                                             ? (unsigned)rintf(ldexpf(fabsf(v[a]), 18 - 64))
                                             : 262143u;
  } else if (m > 0.0f) {
    int e;
    frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    if (e >= -62) {
      if (e > 64) e = 64;
      float big = 0.0f;
#pragma unroll
      for (int a = 0; a < D; a++) big = fmaxf(big, rintf(ldexpf(fabsf(v[a]), 18 - e)));
      if (big >= 262144.0f && e < 64) e++;  // rounding reached 2^18
#pragma unroll
      for (int a = 0; a < D; a++)
        mag[a] = (unsigned)fminf(rintf(ldexpf(fabsf(v[a]), 18 - e)), 262143.0f);
      E = e + 63;
    }
  }
  uint64_t w = (uint64_t)E;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const uint64_t s = signbit(v[a]) && mag[a] ? 1ull : 0ull;
    w |= ((uint64_t)mag[a] | (s << 18)) << (7 + 19 * a);
  }
  return w;
}

template <int D>
__device__ __forceinline__ void quant_v_decode(uint64_t w, float v[3]) {
  const int E = (int)(w & 127u);
  const float scale = E ? ldexpf(1.0f, E - 63 - 18) : 0.0f;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const unsigned f = (unsigned)(w >> (7 + 19 * a)) & 0x7FFFFu;
    const float m = (float)(f & 0x3FFFFu) * scale;
    v[a] = (f >> 18) ? -m : m;
  }
}

// Accessors.  Q: packed layout; G: use_g2p2g layout (packed only: no C).
// Stores take the value by reference and leave in it what a later load
// returns (the reference reads its quantised fields back, e.g.
// x += dt * v reads the stored v, mpm_solver.py:724).
template <int D, bool Q, bool G = false>
struct Particles;

template <int D, bool G>
struct Particles<D, false, G> {
  using L = Layout<D>;
  __device__ static void load_x(const float* pf, int64_t cap, int p, float x[3]) {
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = pf[(int64_t)(L::X + a) * cap + p];
  }
  __device__ static void store_x(float* pf, int64_t cap, int p, float x[3]) {
#pragma unroll
    for (int a = 0; a < D; a++) pf[(int64_t)(L::X + a) * cap + p] = x[a];
  }
  __device__ static void load_v(const float* pf, int64_t cap, int p, float v[3]) {
#pragma unroll
    for (int a = 0; a < D; a++) v[a] = pf[(int64_t)(L::V + a) * cap + p];
  }
  __device__ static void store_v(float* pf, int64_t cap, int p, float v[3]) {
#pragma unroll
    for (int a = 0; a < D; a++) pf[(int64_t)(L::V + a) * cap + p] = v[a];
  }
  __device__ static void load_C(const float* pf, int64_t cap, int p, float C[3][3]) {
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) C[a][e] = pf[(int64_t)(L::C + a * D + e) * cap + p];
  }
  __device__ static void store_C(float* pf, int64_t cap, int p, const float C[3][3]) {
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) pf[(int64_t)(L::C + a * D + e) * cap + p] = C[a][e];
  }
  __device__ static void load_F(const float* pf, int64_t cap, int p, float F[3][3]) {
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) F[a][e] = pf[(int64_t)(L::F + a * D + e) * cap + p];
  }
  __device__ static void store_F(float* pf, int64_t cap, int p, float F[3][3]) {
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) pf[(int64_t)(L::F + a * D + e) * cap + p] = F[a][e];
  }
  __device__ static float load_Jp(const float* pf, int64_t cap, int p) {
    return pf[(int64_t)L::JP * cap + p];
  }
  __device__ static void store_Jp(float* pf, int64_t cap, int p, float Jp) {
    pf[(int64_t)L::JP * cap + p] = Jp;
  }
};

template <int D, bool G>
struct Particles<D, true, G> {
  using K = PackedLayout<D>;
  __device__ static uint64_t ld64(const float* pf, int64_t cap, int w, int p) {
    const unsigned* u = reinterpret_cast<const unsigned*>(pf);
    return (uint64_t)u[(int64_t)w * cap + p] | ((uint64_t)u[(int64_t)(w + 1) * cap + p] << 32);
  }
  __device__ static void st64(float* pf, int64_t cap, int w, int p, uint64_t v) {
    unsigned* u = reinterpret_cast<unsigned*>(pf);
    u[(int64_t)w * cap + p] = (unsigned)v;
    u[(int64_t)(w + 1) * cap + p] = (unsigned)(v >> 32);
  }
  __device__ static void load_x(const float* pf, int64_t cap, int p, float x[3]) {
    const uint64_t w = ld64(pf, cap, K::X, p);
#pragma unroll
    for (int a = 0; a < D; a++) {
      const int c = (int)((unsigned)(w >> (21 * a)) & 0x1FFFFFu);
      x[a] = quant_x_value((c << 11) >> 11);
    }
  }
  __device__ static void store_x(float* pf, int64_t cap, int p, float x[3]) {
    uint64_t w = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
      const int c = quant_x_code(x[a]);
      x[a] = quant_x_value(c);
      w |= (uint64_t)((unsigned)c & 0x1FFFFFu) << (21 * a);
    }
    st64(pf, cap, K::X, p, w);
  }
  __device__ static void load_v(const float* pf, int64_t cap, int p, float v[3]) {
    quant_v_decode<D>(ld64(pf, cap, K::V, p), v);
  }
  __device__ static void store_v(float* pf, int64_t cap, int p, float v[3]) {
    const uint64_t w = quant_v_encode<D>(v);
    quant_v_decode<D>(w, v);
    st64(pf, cap, K::V, p, w);
  }
  __device__ static void load_C(const float* pf, int64_t cap, int p, float C[3][3]) {
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) C[a][e] = G ? 0.0f : pf[(int64_t)(K::C + a * D + e) * cap + p];
  }
  __device__ static void store_C(float* pf, int64_t cap, int p, const float C[3][3]) {
    if (G) return;
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) pf[(int64_t)(K::C + a * D + e) * cap + p] = C[a][e];
  }
  __device__ static void load_F(const float* pf, int64_t cap, int p, float F[3][3]) {
    const unsigned* u = reinterpret_cast<const unsigned*>(pf);
#pragma unroll
    for (int k = 0; k < K::NFW; k++) {
      const unsigned w = u[(int64_t)(K::F + k) * cap + p];
      const int i0 = 2 * k, i1 = 2 * k + 1;
      F[i0 / D][i0 % D] = quant_f_value((int)(short)(w & 0xFFFFu));
      if (i1 < D * D) F[i1 / D][i1 % D] = quant_f_value((int)(short)(w >> 16));
    }
  }
  __device__ static void store_F(float* pf, int64_t cap, int p, float F[3][3]) {
    unsigned* u = reinterpret_cast<unsigned*>(pf);
#pragma unroll
    for (int k = 0; k < K::NFW; k++) {
      const int i0 = 2 * k, i1 = 2 * k + 1;
      const int c0 = quant_f_code(F[i0 / D][i0 % D]);
      F[i0 / D][i0 % D] = quant_f_value(c0);
      unsigned w = (unsigned)c0 & 0xFFFFu;
      if (i1 < D * D) {
        const int c1 = quant_f_code(F[i1 / D][i1 % D]);
        F[i1 / D][i1 % D] = quant_f_value(c1);
        w |= ((unsigned)c1 & 0xFFFFu) << 16;
      }
      u[(int64_t)(K::F + k) * cap + p] = w;
    }
  }
  __device__ static float load_Jp(const float* pf, int64_t cap, int p) {
    return pf[(int64_t)K::jp(G) * cap + p];
  }
  __device__ static void store_Jp(float* pf, int64_t cap, int p, float Jp) {
    pf[(int64_t)K::jp(G) * cap + p] = Jp;
  }
};

}  // namespace mpm
