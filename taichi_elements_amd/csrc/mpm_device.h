// Device-side building blocks of the MLS-MPM substep for gfx950:
// quadratic B-spline stencil set-up, 2x2 / 3x3 SVD, the four constitutive
// models of p2g and the grid post-process operators.  Every function cites
// the reference lines of engine/mpm_solver.py that it computes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_particles.h"

namespace mpm {

enum Material : int { kWater = 0, kElastic = 1, kSnow = 2, kSand = 3, kStationary = 4 };
enum PostKind : int { kPostBBox = 0, kPostPlane = 1, kPostSphere = 2 };
enum Surface : int { kSticky = 0, kSlip = 1, kSeparate = 2 };

constexpr int kMaxPost = 16;
constexpr int kBlock = 4;  // grid block edge (nodes); 4^3 in 3D, 4^2 in 2D
constexpr int kTile = kBlock + 2;  // nodes touched by the particles of one block

template <int D>
struct Geo;
template <>
struct Geo<3> {
  static constexpr int kNodesPerBlock = 64;
  static constexpr int kTileNodes = 216;
  static constexpr int kStencil = 27;
};
template <>
struct Geo<2> {
  static constexpr int kNodesPerBlock = 16;
  static constexpr int kTileNodes = 36;
  static constexpr int kStencil = 9;
};

struct PostOp {
  int kind, surface;
  float point[3];
  float normal[3];
  float radius, friction;
  // bounding box (mpm_solver.py:600-616): node index I_a < lo[a] stops a
  // negative v_a, I_a >= hi[a] a positive one; [padding, res - padding), or
  // [-grid_size/2 + padding, grid_size/2 - padding) for unbounded
  int lo[3], hi[3];
};

// Constants of one substep, passed by value to every kernel.  Host-side
// Python floats of MPMSolver.__init__ (mpm_solver.py:82-87, 200-210) are
// rounded to f32 here, as the reference's kernels see them.
struct Params {
  float dx, inv_dx, dt, p_mass, mu_0, lambda_0, alpha, water_density;
  float stress_scale;   // (-dt * p_vol) * 4 * inv_dx^2, mpm_solver.py:569
  float c4;             // 4 * inv_dx, mpm_solver.py:721
  float sand_coef;      // (d * lambda_0 + 2 mu_0) / (2 mu_0), mpm_solver.py:336-337
  float gravity[3];
  int res[3];
  int padding;
  int blo[3];           // origin of the scene's block window (in blocks)
  int bnl[3];           // extent of the scene's block window (in blocks)
  int bn[3];            // extent of the combined window of all scenes (in blocks)
  int zoff;             // the scene's first block along the stacking axis (last axis)
  int skip;             // 1: the scene sits this substep out (state and grid untouched)
  // Differentiable mode: per-particle Young's modulus E_p (indexed by slot)
  // and a global Poisson ratio nu.  mu_0 = E_p kmu, lambda_0 = E_p kla with
  // kmu = 1 / (2 (1 + nu)), kla = nu / ((1 + nu)(1 - 2 nu)); dkmu, dkla are
  // their derivatives in nu.  youngs == nullptr: the global mu_0, lambda_0.
  const float* youngs;
  float kmu, kla, dkmu, dkla;
  float dsand_coef;     // d sand_coef / d nu (sand_coef = 1 + d kla / (2 kmu) there)
  // use_g2p2g mode (mpm_solver.py:363-485, 773-787): 1 selects the fused
  // transfer; v_allowed = dx * g2p2g_allowed_cfl / dt clamps every grid
  // velocity component (mpm:589, 596-598; <= 0: no clamp).
  int g2p2g;
  float v_allowed;
  // support_plasticity (mpm_solver.py:62, 135-136): 0 drops Jp -- no
  // hardening, no Jp updates, sand without its plastic branch
  int plastic;
  int n_post;
  PostOp post[kMaxPost];
};

// Workgroup barrier for hand-offs through LDS only: waits for this wave's
// LDS operations, not for its global stores.  __syncthreads() also waits for
// every outstanding global store of the wave (a workgroup-scope release),
// which puts the HBM write latency of the particle and tile stores on the
// critical path of kernels whose stores no other wave of the workgroup reads.
// The fences are workgroup-scope and restricted to LDS ("local"), so the
// compiler orders LDS accesses across the barrier and waits for LDS only.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// base = floor(x * inv_dx - 0.5), fx = x * inv_dx - base (mpm_solver.py:497,
// 503).  Contraction is disabled so the binning pass and both transfers see
// bit-identical bases.
__device__ __forceinline__ void base_fx(float xa, float inv_dx, int& base, float& fx) {
#pragma clang fp contract(off)
  float t = xa * inv_dx;
  float u = t - 0.5f;
  base = (int)floorf(u);
  fx = t - (float)base;
}

// One particle's share of one x line of its stencil, the nodes o = (0..2,
// oy, oz) (p2g, mpm_solver.py:575-584, with the affine term folded as
// q = mass v - Ad fx, Ad = A dx): node ox receives wn[ox] (r[ox], mass) with
// the node weight wn[ox] = wx[ox] (wy[oy] wz[oz]) and r[ox] = q + Ad o,
// stepped along x as b, b + Ad e_x, b + 2 Ad e_x from b = q + Ad (0, oy, oz).
// The caller adds these with one fma per channel straight into the node it
// read from the tile.  oy and oz are compile-time after unrolling.
template <int D>
__device__ __forceinline__ void line_terms(int oy, int oz, const float Ad[3][3],
                                           const float q[3], const float w[3][3],
                                           float wn[3], float r[3][3]) {
  const float wyz = D == 3 ? w[1][oy] * w[2][oz] : w[1][oy];
#pragma unroll
  for (int ox = 0; ox < 3; ox++) wn[ox] = w[0][ox] * wyz;
#pragma unroll
  for (int a = 0; a < D; a++) {
    float b = q[a];
    if (oy == 1) b += Ad[a][1];
    if (oy == 2) b += 2.0f * Ad[a][1];
    if (D == 3 && oz == 1) b += Ad[a][2];
    if (D == 3 && oz == 2) b += 2.0f * Ad[a][2];
    r[0][a] = b;
    r[1][a] = b + Ad[a][0];
    r[2][a] = r[1][a] + Ad[a][0];
  }
}

// Lanes per grid node for the grid operators: as many (a power of two up to
// src) as the launched threads cover in one pass over the active nodes.
__device__ __forceinline__ int node_lanes(int64_t nodes, int src) {
  int lanes = src;
  while (lanes > 1 && nodes * lanes > (int64_t)gridDim.x * blockDim.x) lanes >>= 1;
  return lanes;
}

// Sum of SRC per-source partial sums of one grid node, computed by `lanes`
// aligned lanes (a power of two dividing SRC): lane `sub` sums the sources
// [sub * SRC / lanes, (sub + 1) * SRC / lanes) with part(s), then the lanes
// exchange.  Both steps follow one binary tree over the sources
// ((s0 + s1) + (s2 + s3)) + ..., so every lane count gives bitwise the same
// total (on every lane of the group).
template <int SRC, class Part>
__device__ __forceinline__ float4 tree_sum(int lanes, int sub, Part part) {
  const int per = SRC / lanes;  // uniform: every lane runs the same part() calls
  float4 v[SRC];
#pragma unroll
  for (int k = 0; k < SRC; k++)
    v[k] = k < per ? part(sub * per + k) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int w = 1; w < SRC; w <<= 1)
#pragma unroll
    for (int k = 0; k + w < SRC; k += 2 * w) {
      v[k].x += v[k + w].x;
      v[k].y += v[k + w].y;
      v[k].z += v[k + w].z;
      v[k].w += v[k + w].w;
    }
  float4 r = v[0];
  for (int off = 1; off < lanes; off <<= 1) {
    r.x += __shfl_xor(r.x, off);
    r.y += __shfl_xor(r.y, off);
    r.z += __shfl_xor(r.z, off);
    r.w += __shfl_xor(r.w, off);
  }
  return r;
}

// Quadratic B-spline weights, mpm_solver.py:505.
__device__ __forceinline__ void bspline(float f, float w[3]) {
  w[0] = 0.5f * (1.5f - f) * (1.5f - f);
  w[1] = 0.75f - (f - 1.0f) * (f - 1.0f);
  w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
}

__device__ __forceinline__ int floor_div4(int b) { return b >= 0 ? (b >> 2) : -((-b + 3) >> 2); }

// The cell a lane owns in the cell-major scatter.  3D: lane bits
// (x, x, z, z, y, y), so the 16 lanes of each quarter-wave hold x = 0..3 and
// z = 0..3 at one y: their float4 nodes x + 36 z (+ a common offset) of the
// 6-wide LDS tile are distinct modulo 16 and modulo 8 (36 z = 0, 4, 8, 12
// mod 16), which keeps the 128-bit tile accesses free of bank conflicts
// (the x-fastest order put nodes 0..3 and 6..9 in one group: 2-way
// conflicts).  2D: x fastest, 16 cells per lane group.
template <int D>
__device__ __forceinline__ int scatter_cell_of_lane(int lane) {
  return D == 3 ? (lane & 3) + 4 * ((lane >> 4) + 4 * ((lane >> 2) & 3)) : lane % 16;
}

// Whether a grid node's mass takes part in the normalisation
// (mpm_solver.py:591 tests m > 0).  A mass below the smallest normal float
// comes only from stencil weights of ~1e-15 per axis (a particle within a
// float step of a (k + 1/2) dx plane, with nothing else on the node); 1/m
// overflows there and the node's velocity would be infinite.  The reference
// on GPUs, built with flush-to-zero, reads such a mass as 0 and skips the
// node; so does this engine.
__device__ __forceinline__ bool node_has_mass(float m) { return m >= 1.17549435e-38f; }

// One pass of the cell-major scatter of p2g and g2p2g.  The lane owns cell
// lc of the block (its particles share their 27 (9) stencil nodes) and adds
// the contributions w_o (q + Ad o, mass) of NB of them -- slots k0, k0 +
// SUBS, ... of its cell's ccnt particles, listed from order[cbeg]; an empty
// slot adds nothing -- to its private LDS tile tp ((4+2)^d nodes) with one
// plain read-modify-write per node.  stage[r][pt] holds particle pt's record
// (3D: Ad row-major, q, fx, mass; 2D: Ad, q, fx, mass).
template <int D, int NB, int SUBS, int ROW>
__device__ __forceinline__ void cell_scatter_pass(int k0, int ccnt, int cbeg,
                                                  const unsigned char* order,
                                                  const float4 (*stage)[ROW], const int lc[3],
                                                  float4* tp) {
  constexpr int kT = kBlock + 2;
  float Ad[NB][3][3], q[NB][3], mass[NB], w[NB][3][3];
#pragma unroll
  for (int bi = 0; bi < NB; bi++) {
    const int k = k0 + bi * SUBS;
    float4 s0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), s1 = s0, s2 = s0, s3 = s0;
    if (k < ccnt) {
      const int pt = order[cbeg + k];
      s0 = stage[0][pt];
      s1 = stage[1][pt];
      s2 = stage[2][pt];
      s3 = stage[3][pt];
    }
    float fx[3];
    if (D == 3) {
      Ad[bi][0][0] = s0.x; Ad[bi][0][1] = s0.y; Ad[bi][0][2] = s0.z;
      Ad[bi][1][0] = s0.w; Ad[bi][1][1] = s1.x; Ad[bi][1][2] = s1.y;
      Ad[bi][2][0] = s1.z; Ad[bi][2][1] = s1.w; Ad[bi][2][2] = s2.x;
      q[bi][0] = s2.y; q[bi][1] = s2.z; q[bi][2] = s2.w;
      fx[0] = s3.x; fx[1] = s3.y; fx[2] = s3.z;
      mass[bi] = s3.w;
    } else {
      Ad[bi][0][0] = s0.x; Ad[bi][0][1] = s0.y; Ad[bi][1][0] = s0.z; Ad[bi][1][1] = s0.w;
      Ad[bi][0][2] = Ad[bi][1][2] = Ad[bi][2][0] = Ad[bi][2][1] = Ad[bi][2][2] = 0.0f;
      q[bi][0] = s1.x; q[bi][1] = s1.y; q[bi][2] = 0.0f;
      fx[0] = s1.z; fx[1] = s1.w; fx[2] = 0.0f;
      mass[bi] = s2.x;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) bspline(fx[a], w[bi][a]);
    if (k >= ccnt) {  // an empty slot contributes exactly nothing
#pragma unroll
      for (int a = 0; a < 3; a++) w[bi][a][0] = w[bi][a][1] = w[bi][a][2] = 0.0f;
    }
  }
#pragma unroll
  for (int oz = 0; oz < (D == 3 ? 3 : 1); oz++)
#pragma unroll
    for (int oy = 0; oy < 3; oy++) {
      float wn[NB][3], r[NB][3][3];
#pragma unroll
      for (int bi = 0; bi < NB; bi++) line_terms<D>(oy, oz, Ad[bi], q[bi], w[bi], wn[bi], r[bi]);
#pragma unroll
      for (int ox = 0; ox < 3; ox++) {
        const int node = D == 3 ? (lc[0] + ox) + kT * ((lc[1] + oy) + kT * (lc[2] + oz))
                                : (lc[0] + ox) + kT * (lc[1] + oy);
        float4 acc = tp[node];
#pragma unroll
        for (int bi = 0; bi < NB; bi++) {
          acc.x = fmaf(wn[bi][ox], r[bi][ox][0], acc.x);
          acc.y = fmaf(wn[bi][ox], r[bi][ox][1], acc.y);
          if (D == 3) acc.z = fmaf(wn[bi][ox], r[bi][ox][2], acc.z);
          acc.w = fmaf(wn[bi][ox], mass[bi], acc.w);
        }
        tp[node] = acc;
        // Lanes of a wave update distinct nodes in one instruction, but a
        // later node of one lane can be an earlier node of another: keep
        // every LDS access in program order (LDS serves a wave in order).
        __asm__ volatile("" ::: "memory");
      }
    }
}

// The scatter of one lane's cell: passes of up to two particles per lane;
// a pass in which no lane of the wave has a second particle evaluates one
// slot only (sparse chunks, and the last pass of most dense ones).  The
// lanes still in the loop are the active ones, so the test is wave-uniform.
template <int D, int SUBS, int ROW>
__device__ __forceinline__ void cell_scatter(int sub, int ccnt, int cbeg, const unsigned char* order,
                                             const float4 (*stage)[ROW], const int lc[3],
                                             float4* tp) {
  for (int k0 = sub; k0 < ccnt; k0 += 2 * SUBS) {
    if (__any(k0 + SUBS < ccnt))
      cell_scatter_pass<D, 2, SUBS>(k0, ccnt, cbeg, order, stage, lc, tp);
    else
      cell_scatter_pass<D, 1, SUBS>(k0, ccnt, cbeg, order, stage, lc, tp);
  }
}

// --------------------------------------------------------------------------
// SVD F = U diag(s) V^T, U, V in SO(d), s decreasing, s[d-1] signed.
// Replaces ti.svd (mpm_solver.py:525).  3x3: cyclic Jacobi eigen-analysis
// of F^T F with exact rotation angles in the form of Rutishauser (Numer.
// Math. 9, 1966: t = tan theta from the 2x2 subproblem, the off-diagonal
// entry set to zero); three sweeps leave the off-diagonal part of
// V^T F^T F V at ~1e-7 of its norm, as five sweeps of the approximate
// quaternion rotations of McAdams et al. 2011 (UW-Madison TR1690) did, for
// about half the arithmetic.  The rotations are accumulated in V, then one
// Newton-Schulz step V (3 I - V^T V) / 2 makes V orthonormal to rounding;
// a column sort and a Givens QR of F V give U and the singular values.  On the device the rotation angles use the
// hardware square-root, reciprocal and reciprocal-square-root estimates
// (fast_sqrt, fast_rcp, fast_rsqrt; the host build, used by mpm_host_svd,
// uses sqrtf and divisions); the Givens QR divides by sqrtf on both.
// --------------------------------------------------------------------------
__host__ __device__ __forceinline__ float fast_rsqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

__host__ __device__ __forceinline__ float fast_sqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

__host__ __device__ __forceinline__ float fast_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// One Jacobi rotation of the symmetric S in the (P, Q) plane: t = tan theta
// of the smaller rotation (|theta| <= pi/4) that zeroes S_PQ,
// t = 2 S_PQ sgn(d) / (|d| + sqrt(d^2 + 4 S_PQ^2)), d = S_QQ - S_PP (t = 0
// when S_PQ = d = 0); S <- G^T S G with G e_P = c e_P - s e_Q,
// G e_Q = s e_P + c e_Q, and V <- V G.
template <int P, int Q>
__host__ __device__ __forceinline__ void jacobi_rot(float S[3][3], float V[3][3]) {
  constexpr int R = 3 - P - Q;
  const float spq = S[P][Q];
  const float d = S[Q][Q] - S[P][P];
  const float den = fabsf(d) + fast_sqrt(d * d + 4.0f * spq * spq);
  const float t = den > 0.0f ? (d >= 0.0f ? 2.0f : -2.0f) * spq * fast_rcp(den) : 0.0f;
  const float c = fast_rsqrt(1.0f + t * t), s = t * c;
  S[P][P] -= t * spq;
  S[Q][Q] += t * spq;
  S[P][Q] = S[Q][P] = 0.0f;
  const float srp = S[R][P], srq = S[R][Q];
  S[R][P] = S[P][R] = c * srp - s * srq;
  S[R][Q] = S[Q][R] = s * srp + c * srq;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float vp = V[k][P], vq = V[k][Q];
    V[k][P] = c * vp - s * vq;
    V[k][Q] = s * vp + c * vq;
  }
}

template <int I, int J>
__host__ __device__ __forceinline__ void sort_cols(float B[3][3], float V[3][3]) {
  float ni = B[0][I] * B[0][I] + B[1][I] * B[1][I] + B[2][I] * B[2][I];
  float nj = B[0][J] * B[0][J] + B[1][J] * B[1][J] + B[2][J] * B[2][J];
  if (ni < nj) {
#pragma unroll
    for (int r = 0; r < 3; r++) {
      float tb = B[r][I];
      B[r][I] = B[r][J];
      B[r][J] = -tb;
      float tv = V[r][I];
      V[r][I] = V[r][J];
      V[r][J] = -tv;
    }
  }
}

template <int P, int Q, int D>
__host__ __device__ __forceinline__ void givens_qr(float B[3][3], float U[3][3]) {
  float a = B[P][P], b = B[Q][P];
  float r2 = a * a + b * b;
  float c = 1.0f, s = 0.0f;
  if (r2 > 1e-30f) {
    // a correctly rounded square root, then a correctly rounded division,
    // not the hardware estimate: that estimate's last-bit error leaves
    // c^2 + s^2 off 1 in a biased way, so U and R = U V^T shrink or grow by
    // ~1e-7, which 2 mu (F - R) F^T turns into stress at F ~ I (2D snow
    // falling freely drifted 4e-4 from the float64 oracle with it).  Nor a
    // correctly rounded 1 / sqrt (estimate plus a Newton step with an exact
    // residual; 2D snow: 1.5e-4): with b = 0, sqrtf(fl(a^2)) returns |a|
    // for nearly every a and a * fl(1 / |a|) rounds to +-1, so c stays at
    // +-1; a correctly rounded 1 / sqrt(fl(a^2)) does not keep that.
    float ir = 1.0f / sqrtf(r2);
    c = a * ir;
    s = b * ir;
  }
#pragma unroll
  for (int k = 0; k < D; k++) {
    float bp = B[P][k], bq = B[Q][k];
    B[P][k] = c * bp + s * bq;
    B[Q][k] = -s * bp + c * bq;
    float up = U[k][P], uq = U[k][Q];
    U[k][P] = c * up + s * uq;
    U[k][Q] = -s * up + c * uq;
  }
}

__host__ __device__ __forceinline__ void svd3(const float A[3][3], float U[3][3], float sig[3],
                                              float V[3][3]) {
  float S[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      S[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
      U[i][j] = i == j ? 1.0f : 0.0f;
      V[i][j] = i == j ? 1.0f : 0.0f;
    }
#pragma unroll
  for (int sweep = 0; sweep < 3; sweep++) {
    jacobi_rot<0, 1>(S, V);
    jacobi_rot<1, 2>(S, V);
    jacobi_rot<0, 2>(S, V);
  }
  {
    // V <- V (3 I - V^T V) / 2: the rounding of the accumulated rotations
    // leaves V^T V = I + O(1e-7); one Newton-Schulz step squares that
    float M[3][3], W[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = i; j < 3; j++) {
        const float g = V[0][i] * V[0][j] + V[1][i] * V[1][j] + V[2][i] * V[2][j];
        M[i][j] = M[j][i] = (i == j ? 1.5f : 0.0f) - 0.5f * g;
      }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) W[i][j] = V[i][0] * M[0][j] + V[i][1] * M[1][j] + V[i][2] * M[2][j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) V[i][j] = W[i][j];
  }
  float B[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) B[i][j] = A[i][0] * V[0][j] + A[i][1] * V[1][j] + A[i][2] * V[2][j];
  sort_cols<0, 1>(B, V);
  sort_cols<0, 2>(B, V);
  sort_cols<1, 2>(B, V);
  givens_qr<0, 1, 3>(B, U);
  givens_qr<0, 2, 3>(B, U);
  givens_qr<1, 2, 3>(B, U);
  sig[0] = B[0][0];
  sig[1] = B[1][1];
  sig[2] = B[2][2];
}

// 2x2: one exact Jacobi rotation of F^T F, column sort, one Givens QR.
__host__ __device__ __forceinline__ void svd2(const float A[3][3], float U[3][3], float sig[3], float V[3][3]) {
  float s00 = A[0][0] * A[0][0] + A[1][0] * A[1][0];
  float s11 = A[0][1] * A[0][1] + A[1][1] * A[1][1];
  float s01 = A[0][0] * A[0][1] + A[1][0] * A[1][1];
  float c = 1.0f, s = 0.0f;
  if (s01 != 0.0f) {
    float tau = (s00 - s11) / (2.0f * s01);
    float t = (tau >= 0.0f ? 1.0f : -1.0f) / (fabsf(tau) + sqrtf(1.0f + tau * tau));
    c = 1.0f / sqrtf(1.0f + t * t);
    s = t * c;
  }
  V[0][0] = c;
  V[1][0] = s;
  V[0][1] = -s;
  V[1][1] = c;
  float B[3][3];
  B[0][0] = A[0][0] * c + A[0][1] * s;
  B[1][0] = A[1][0] * c + A[1][1] * s;
  B[0][1] = -A[0][0] * s + A[0][1] * c;
  B[1][1] = -A[1][0] * s + A[1][1] * c;
  float n0 = B[0][0] * B[0][0] + B[1][0] * B[1][0];
  float n1 = B[0][1] * B[0][1] + B[1][1] * B[1][1];
  if (n0 < n1) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
      float tb = B[r][0];
      B[r][0] = B[r][1];
      B[r][1] = -tb;
      float tv = V[r][0];
      V[r][0] = V[r][1];
      V[r][1] = -tv;
    }
  }
  U[0][0] = 1.0f;
  U[0][1] = 0.0f;
  U[1][0] = 0.0f;
  U[1][1] = 1.0f;
  givens_qr<0, 1, 2>(B, U);
  sig[0] = B[0][0];
  sig[1] = B[1][1];
}

// Drucker-Prager return map on the singular values, mpm_solver.py:321-342.
template <int D>
__device__ __forceinline__ void sand_projection(float sig[3], float& Jp, const Params& P) {
  float eps[3], tr = 0.0f;
#pragma unroll
  for (int i = 0; i < D; i++) {
    eps[i] = logf(fmaxf(fabsf(sig[i]), 1e-4f));
    tr += eps[i];
  }
  tr += Jp;
  float hat[3], n2 = 0.0f;
#pragma unroll
  for (int i = 0; i < D; i++) {
    hat[i] = eps[i] - tr / (float)D;
    n2 += hat[i] * hat[i];
  }
  float hat_norm = sqrtf(n2) + 1e-20f;
  if (tr >= 0.0f) {
    Jp = tr;
#pragma unroll
    for (int i = 0; i < D; i++) sig[i] = 1.0f;
  } else {
    Jp = 0.0f;
    float dg = hat_norm + P.sand_coef * tr * P.alpha;
    float k = fmaxf(dg, 0.0f) / hat_norm;
#pragma unroll
    for (int i = 0; i < D; i++) sig[i] = expf(eps[i] - k * hat[i]);
  }
}

// F rounded to the packed layout's 16-bit fixed point (what a store then a
// load of the quantised F field returns).
template <int D>
__device__ __forceinline__ void quant_f_round(float F[3][3]) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) F[i][j] = quant_f_value(quant_f_code(F[i][j]));
}

// Per-particle F / Jp update and Kirchhoff stress, mpm_solver.py:507-568.
// F is updated in place; stress receives tau (before the p2g scale).
// G2P2G = true: the update inside g2p2g (mpm_solver.py:416-473), which
// differs for water: F is not rebuilt from Jp first, the hardening
// exp(10 (1 - Jp)) applies, and Jp is left unchanged (the volume is kept
// in F[0][0] only).  QF (quantised storage, g2p2g only): the new F is
// clamped to +-F_bound and, like every F the reference writes to its
// quantised field and reads back before the stress, rounded to the packed
// format (mpm_solver.py:413-416, 428, 444-456).  P.plastic = 0
// (support_plasticity=False): h = 1, Jp is neither read nor updated, water
// restarts from F = I, sand keeps F and has no stress.
template <int D, bool G2P2G = false, bool QF = false>
__device__ __forceinline__ void constitutive(int mat, const Params& P, float mu_0, float lambda_0,
                                             float F[3][3], const float C[3][3], float& Jp,
                                             float stress[3][3]) {
  const bool plastic = P.plastic != 0;
  float G[3][3];
  if (mat == kWater && !G2P2G) {  // mpm_solver.py:508-511
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) F[i][j] = i == j ? 1.0f : 0.0f;
    if (plastic) F[0][0] = Jp;
  }
#pragma unroll
  for (int i = 0; i < D; i++)  // (I + dt C) F, mpm_solver.py:513
#pragma unroll
    for (int j = 0; j < D; j++) {
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < D; k++) s += ((i == k ? 1.0f : 0.0f) + P.dt * C[i][k]) * F[k][j];
      G[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) F[i][j] = QF ? fmaxf(-kFBound, fminf(kFBound, G[i][j])) : G[i][j];
  if (QF) quant_f_round<D>(F);

  if (mat == kWater) {
    // Water (mpm_solver.py:515-551 with mu = 0, h = 1): the singular values
    // are never clamped, so J = prod(sigma) = det F and the deviatoric term
    // vanishes; F = diag(J, 1, ..), Jp = J, tau = lambda_0 J (J - 1) I.
    const float J = D == 3 ? F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
                                 F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                                 F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0])
                           : F[0][0] * F[1][1] - F[0][1] * F[1][0];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        F[i][j] = i == j ? 1.0f : 0.0f;
        stress[i][j] = 0.0f;
      }
    F[0][0] = J;
    float la = lambda_0;
    if (G2P2G) {
      if (plastic) la = lambda_0 * __expf(10.0f * (1.0f - Jp));  // mpm:419-424 (Jp stays 1 unless set)
    } else if (plastic) {
      Jp = J;
    }
    const float iso = la * J * (J - 1.0f);
#pragma unroll
    for (int i = 0; i < D; i++) stress[i][i] = iso;
    return;
  }

  // Hardening, mpm_solver.py:515-524 (p2g: every material but water) and
  // :419-424 (g2p2g: every material): h = exp(10 (1 - Jp)) for snow and for
  // stationary particles, whose Jp a caller may set; elastic is fixed at
  // 0.3, and sand's stress uses mu_0, lambda_0, so its h is never formed.
  float h = 1.0f;
  if (plastic && (mat == kSnow || mat == kStationary)) h = __expf(10.0f * (1.0f - Jp));
  if (mat == kElastic) h = 0.3f;
  float mu = mu_0 * h, la = lambda_0 * h;

  if (mat == kSand && !plastic) {  // mpm_solver.py:552-553: no stress, F kept
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) stress[i][j] = 0.0f;
    return;
  }

  float U[3][3], V[3][3], sig[3];
  if (D == 3)
    svd3(F, U, sig, V);
  else
    svd2(F, U, sig, V);

  float J = 1.0f;
  if (mat != kSand) {  // mpm_solver.py:527-536
#pragma unroll
    for (int i = 0; i < D; i++) {
      float ns = sig[i];
      if (mat == kSnow) ns = fminf(fmaxf(sig[i], 1.0f - 2.5e-2f), 1.0f + 4.5e-3f);
      if (plastic) {
        if (mat == kSnow)
          Jp *= sig[i] / ns;
        else if (!(fabsf(sig[i]) > 0.0f && fabsf(sig[i]) < INFINITY))
          Jp = NAN;  // sig / sig: exactly 1 unless sig is 0, infinite or NaN
      }
      sig[i] = ns;
      J *= ns;
    }
  }
  if (mat == kSand) sand_projection<D>(sig, Jp, P);  // mpm_solver.py:554
  if (mat == kSnow || mat == kSand) {                // F = U sig V^T, mpm_solver.py:545, 555
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) s += U[i][k] * sig[k] * V[j][k];
        F[i][j] = s;
      }
    if (QF) quant_f_round<D>(F);
  }
  if (mat != kSand) {  // tau = 2 mu (F - U V^T) F^T + la J (J - 1) I, mpm_solver.py:549-551
    float Dm[3][3];
    const float two_mu = 2.0f * mu;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        float r = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) r += U[i][k] * V[j][k];
        Dm[i][j] = two_mu * (F[i][j] - r);
      }
    const float iso = la * J * (J - 1.0f);
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) s += Dm[i][k] * F[j][k];
        stress[i][j] = s + (i == j ? iso : 0.0f);
      }
  } else {  // tau = U diag(2 mu ln s / s + la sum ln s / s) V^T F^T, mpm_solver.py:556-566
    float lsum = 0.0f, center[3];
#pragma unroll
    for (int i = 0; i < D; i++) {
      float l = logf(sig[i]);
      lsum += l;
      center[i] = 2.0f * mu_0 * l * (1.0f / sig[i]);
    }
#pragma unroll
    for (int i = 0; i < D; i++) center[i] += lambda_0 * lsum * (1.0f / sig[i]);
    float T[3][3];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) s += U[i][k] * center[k] * V[j][k];
        T[i][j] = s;
      }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) s += T[i][k] * F[j][k];
        stress[i][j] = s;
      }
  }
}

// One grid_postprocess operator on one node velocity (mpm_solver.py:600-687).
template <int D>
__device__ __forceinline__ void apply_post_op(const PostOp& op, const Params& P, const int I[3],
                                              float v[3]) {
  if (op.kind == kPostBBox) {  // mpm_solver.py:613-616
#pragma unroll
    for (int a = 0; a < D; a++) {
      if (I[a] < op.lo[a] && v[a] < 0.0f) v[a] = 0.0f;
      if (I[a] >= op.hi[a] && v[a] > 0.0f) v[a] = 0.0f;
    }
  } else if (op.kind == kPostPlane) {  // mpm_solver.py:660-685
    float dot = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) dot += ((float)I[a] * P.dx - op.point[a]) * op.normal[a];
    if (dot < 0.0f) {
      if (op.surface == kSticky) {
#pragma unroll
        for (int a = 0; a < D; a++) v[a] = 0.0f;
      } else {
        float nc = 0.0f;
#pragma unroll
        for (int a = 0; a < D; a++) nc += op.normal[a] * v[a];
        float proj = op.surface == kSlip ? nc : fminf(nc, 0.0f);
        float vn2 = 0.0f;
#pragma unroll
        for (int a = 0; a < D; a++) {
          v[a] = v[a] - op.normal[a] * proj;
          vn2 += v[a] * v[a];
        }
        float vn = sqrtf(vn2);
        if (nc < 0.0f && vn > 1e-30f) {
          float mag = fmaxf(0.0f, vn + nc * op.friction);
          float inv = 1.0f / vn;
#pragma unroll
          for (int a = 0; a < D; a++) v[a] = v[a] * inv * mag;
        }
      }
    }
  } else if (op.kind == kPostSphere) {  // mpm_solver.py:621-640
    float off[3], n2 = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) {
      off[a] = (float)I[a] * P.dx - op.point[a];
      n2 += off[a] * off[a];
    }
    if (n2 < op.radius * op.radius) {
      if (op.surface == kSticky) {
#pragma unroll
        for (int a = 0; a < D; a++) v[a] = 0.0f;
      } else {
        float inv = 1.0f / (sqrtf(n2) + 1e-5f);
        float nc = 0.0f;
#pragma unroll
        for (int a = 0; a < D; a++) nc += off[a] * inv * v[a];
        float proj = op.surface == kSlip ? nc : fminf(nc, 0.0f);
#pragma unroll
        for (int a = 0; a < D; a++) v[a] = v[a] - off[a] * inv * proj;
      }
    }
  }
}

// grid_postprocess operators on one node velocity, in the order added.
template <int D>
__device__ __forceinline__ void apply_post(const Params& P, const int I[3], float v[3]) {
  for (int k = 0; k < P.n_post; k++) apply_post_op<D>(P.post[k], P, I, v);
}

// Adjoint of one operator: vbar (the adjoint of its output) becomes the
// adjoint of its input v.
//  - walls and sticky surfaces zero what they zero (a mask);
//  - slip / separate planes and spheres project out the normal part:
//    u = v - n p(n . v) with p = id (slip) or min(., 0) (separate), so
//    vbar <- ubar - n p'(n . v) (n . ubar);
//  - friction on a plane, out = u max(0, |u| + f (n . v)) / |u| when
//    n . v < 0: ubar = (1 + f nc / |u|) vbar - f nc (u . vbar) u / |u|^3 and
//    the normal component nc = n . v receives f (u . vbar) / |u| (zero where
//    the max clamps the velocity to rest).
template <int D>
__device__ __forceinline__ void post_op_adjoint(const PostOp& op, const Params& P, const int I[3],
                                                const float v[3], float vb[3]) {
  if (op.kind == kPostBBox) {
#pragma unroll
    for (int a = 0; a < D; a++)
      if ((I[a] < op.lo[a] && v[a] < 0.0f) || (I[a] >= op.hi[a] && v[a] > 0.0f)) vb[a] = 0.0f;
    return;
  }
  float nrm[3] = {0.0f, 0.0f, 0.0f};
  bool inside;
  if (op.kind == kPostPlane) {
    float dot = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) {
      dot += ((float)I[a] * P.dx - op.point[a]) * op.normal[a];
      nrm[a] = op.normal[a];
    }
    inside = dot < 0.0f;
  } else {
    float off[3], n2 = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) {
      off[a] = (float)I[a] * P.dx - op.point[a];
      n2 += off[a] * off[a];
    }
    inside = n2 < op.radius * op.radius;
    const float inv = 1.0f / (sqrtf(n2) + 1e-5f);
#pragma unroll
    for (int a = 0; a < D; a++) nrm[a] = off[a] * inv;
  }
  if (!inside) return;
  if (op.surface == kSticky) {
#pragma unroll
    for (int a = 0; a < D; a++) vb[a] = 0.0f;
    return;
  }
  float nc = 0.0f;
#pragma unroll
  for (int a = 0; a < D; a++) nc += nrm[a] * v[a];
  const float dproj = (op.surface == kSlip || nc < 0.0f) ? 1.0f : 0.0f;
  const float proj = op.surface == kSlip ? nc : fminf(nc, 0.0f);
  float ub[3] = {vb[0], vb[1], vb[2]}, ncb = 0.0f;
  if (op.kind == kPostPlane) {
    float u[3] = {0.0f, 0.0f, 0.0f}, un2 = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) {
      u[a] = v[a] - nrm[a] * proj;
      un2 += u[a] * u[a];
    }
    const float un = sqrtf(un2);
    if (nc < 0.0f && un > 1e-30f) {
      const float mag = un + nc * op.friction;
      if (mag <= 0.0f) {
#pragma unroll
        for (int a = 0; a < D; a++) vb[a] = 0.0f;
        return;
      }
      float uv = 0.0f;
#pragma unroll
      for (int a = 0; a < D; a++) uv += u[a] * vb[a];
      const float g = nc * op.friction / un;
#pragma unroll
      for (int a = 0; a < D; a++) ub[a] = (1.0f + g) * vb[a] - g * uv * u[a] / un2;
      ncb = op.friction * uv / un;
    }
  }
  float nu = 0.0f;
#pragma unroll
  for (int a = 0; a < D; a++) nu += nrm[a] * ub[a];
#pragma unroll
  for (int a = 0; a < D; a++) vb[a] = ub[a] - nrm[a] * dproj * nu + nrm[a] * ncb;
}

// Adjoint of the whole post-process list at a node whose velocity entering
// it is v0: the operators in reverse order, each at its own input (the
// forward re-run up to it).
template <int D>
__device__ __forceinline__ void apply_post_adjoint(const Params& P, const int I[3],
                                                   const float v0[3], float vb[3]) {
  for (int k = P.n_post - 1; k >= 0; k--) {
    float v[3] = {v0[0], v0[1], v0[2]};
    for (int j = 0; j < k; j++) apply_post_op<D>(P.post[j], P, I, v);
    post_op_adjoint<D>(P.post[k], P, I, v, vb);
  }
}

// ---------------------------------------------------------------------------
// combined block window: block ids <-> coordinates, scenes
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void decode_local(int bid, const Params& P, int lb[3]) {
  lb[0] = bid % P.bn[0];
  int r = bid / P.bn[0];
  if (D == 3) {
    lb[1] = r % P.bn[1];
    lb[2] = r / P.bn[1];
  } else {
    lb[1] = r;
    lb[2] = 0;
  }
}

template <int D>
__device__ __forceinline__ int encode_local(const int lb[3], const Params& P) {
  return D == 3 ? (lb[2] * P.bn[1] + lb[1]) * P.bn[0] + lb[0] : lb[1] * P.bn[0] + lb[0];
}

// Scenes are stacked along the last axis of one combined block window; the
// slab table maps a combined block coordinate on that axis to its scene.
template <int D>
__device__ __forceinline__ int scene_of(const int lb[3], const int* __restrict__ slab) {
  return slab[lb[D - 1]];
}

// scene-local block coordinates of a combined block
template <int D>
__device__ __forceinline__ void to_local(const int lb[3], const Params& P, int out[3]) {
  out[0] = lb[0];
  out[1] = lb[1];
  out[2] = lb[2];
  out[D - 1] -= P.zoff;
}

}  // namespace mpm
