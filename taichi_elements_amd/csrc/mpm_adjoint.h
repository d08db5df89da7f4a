// Reverse-mode adjoint of one substep for water and fixed-corotated elastic
// particles (mpm_solver.py:487-584 p2g, 694-724 g2p), device code included
// by mpm_kernels.hip.
//
// Notation per particle: K = I + dt C, F' = K F (water: F = diag(Jp, 1..)),
// F' = U diag(sig) V^T, R = U V^T, J = prod(sig) = det F',
//   tau = 2 mu (F' - R) F'^T + la J (J - 1) I      (water: mu = 0),
//   A = s tau + m C  with s = -dt p_vol 4 inv_dx^2,
// node i: d_i = i - fx (cells), w_i, grad w_i = inv_dx dw/dfx,
//   p2g:  p_i += w_i (m v + A d_i dx),  m_i += w_i m
//   g2p:  v' = sum w_i v_i, C' = c4 sum w_i v_i d_i^T, x' = x + dt v'.
// Given the adjoints of the new state (x', v', C', F', Jp') and of the grid
// (pbar_i, mbar_i, from the grid-operator adjoint), the adjoints of the old
// state are
//   xbar = x'bar + sum_i [ (a . v_i + c4 v_i^T C'bar d_i) + pbar_i . (m v + A d_i dx)
//                          + m mbar_i ] grad w_i  - c4 inv_dx C'bar^T v' - A^T S
//   with a = v'bar + dt x'bar, S = sum_i w_i pbar_i;
//   vbar = m S;  Abar = dx sum_i w_i pbar_i d_i^T;  Cbar = m Abar + (through K);
//   taubar = s Abar, and through tau:
//     mubar = 2 <taubar, (F' - R) F'^T>, labar = tr(taubar) J (J - 1),
//     Jbar = la tr(taubar) (2 J - 1),
//     F'bar = F'bar_out + 2 mu (taubar F' + taubar^T (F' - R)) + Jbar cof(F')
//             + polar adjoint of Rbar = -2 mu taubar F':
//               U Mbar V^T with Mbar_ij = (B_ij - B_ji) / (sig_i + sig_j),
//               B = U^T Rbar V (the derivative of R = U V^T);
//   elastic: Fbar = K^T F'bar, Cbar += dt F'bar F^T, Jpbar = Jp'bar;
//   water:   J = Jp det K, Jbar += F'bar_out[0][0] + Jp'bar,
//            Jpbar = Jbar det K, Cbar += Jbar Jp dt cof(K), Fbar = 0.
// Material parameters: mu_0 = E kmu(nu), lambda_0 = E kla(nu), times
// h = 0.3 for elastic (mpm_solver.py:519-524), giving Ebar and nubar.
#pragma once

#include "mpm_device.h"

namespace mpm {

template <int D>
__device__ __forceinline__ float det_d(const float M[3][3]) {
  return D == 3 ? M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                      M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                      M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])
                : M[0][0] * M[1][1] - M[0][1] * M[1][0];
}

// cofactor matrix = d det(M) / dM
template <int D>
__device__ __forceinline__ void cofactor_d(const float M[3][3], float Cf[3][3]) {
  if (D == 3) {
    Cf[0][0] = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    Cf[0][1] = M[1][2] * M[2][0] - M[1][0] * M[2][2];
    Cf[0][2] = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    Cf[1][0] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
    Cf[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
    Cf[1][2] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
    Cf[2][0] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
    Cf[2][1] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
    Cf[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
  } else {
    Cf[0][0] = M[1][1];
    Cf[0][1] = -M[1][0];
    Cf[1][0] = -M[0][1];
    Cf[1][1] = M[0][0];
  }
}

// Quadratic B-spline derivative dw/dfx (of bspline in mpm_device.h).
__device__ __forceinline__ void bspline_grad(float f, float dw[3]) {
  dw[0] = f - 1.5f;
  dw[1] = -2.0f * (f - 1.0f);
  dw[2] = f - 0.5f;
}

// Per-particle recompute of the constitutive update of p2g for water and
// elastic particles (the forward in mpm_device.h: constitutive): F' = K F,
// its SVD and rotation R, J and the affine matrix A = s tau + m C.
template <int D>
struct Recompute {
  float K[3][3], Fp[3][3], U[3][3], V[3][3], sig[3], R[3][3], J, mu, la, mass, E, h;
  bool water;
};

template <int D>
__device__ __forceinline__ void recompute(const Params& P, int mat, const float C[3][3],
                                          const float F[3][3], float Jp, float E,
                                          Recompute<D>& r) {
  r.water = mat == kWater;
  r.E = E;
  r.h = r.water ? 1.0f : 0.3f;
  r.mu = r.water ? 0.0f : r.h * (P.youngs ? E * P.kmu : P.mu_0);
  r.la = r.h * (P.youngs ? E * P.kla : P.lambda_0);
  r.mass = P.p_mass;
  if (r.water) r.mass *= P.water_density;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    r.sig[a] = 1.0f;
#pragma unroll
    for (int e = 0; e < 3; e++) {
      r.K[a][e] = r.Fp[a][e] = r.U[a][e] = r.V[a][e] = r.R[a][e] = 0.0f;
    }
  }
#pragma unroll
  for (int a = 0; a < D; a++)
#pragma unroll
    for (int e = 0; e < D; e++) r.K[a][e] = (a == e ? 1.0f : 0.0f) + P.dt * C[a][e];
#pragma unroll
  for (int a = 0; a < D; a++)
#pragma unroll
    for (int e = 0; e < D; e++) {
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < D; k++)
        s += r.K[a][k] * (r.water ? (k == e ? (k == 0 ? Jp : 1.0f) : 0.0f) : F[k][e]);
      r.Fp[a][e] = s;
    }
  if (r.water) {
    r.J = det_d<D>(r.Fp);
  } else {
    if (D == 3)
      svd3(r.Fp, r.U, r.sig, r.V);
    else
      svd2(r.Fp, r.U, r.sig, r.V);
    r.J = 1.0f;
#pragma unroll
    for (int a = 0; a < D; a++) r.J *= r.sig[a];
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) s += r.U[a][k] * r.V[e][k];
        r.R[a][e] = s;
      }
  }
}

template <int D>
__device__ __forceinline__ void load_cf(const float* __restrict__ pf, int64_t cap, int p,
                                        float C[3][3], float F[3][3]) {
  using L = Layout<D>;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int e = 0; e < 3; e++) C[a][e] = F[a][e] = 0.0f;
#pragma unroll
  for (int a = 0; a < D; a++)
#pragma unroll
    for (int e = 0; e < D; e++) {
      C[a][e] = pf[(int64_t)(L::C + a * D + e) * cap + p];
      F[a][e] = pf[(int64_t)(L::F + a * D + e) * cap + p];
    }
}

// Step 1 (per particle): the affine matrix A of every binned particle, into
// abuf (D*D rows of cap).
template <int D>
__global__ __launch_bounds__(kThreads) void k_adj_affine(
    const Params* __restrict__ scenes, const int* __restrict__ slab,
    const float* __restrict__ pf, const int* __restrict__ pi, int64_t cap,
    const int* __restrict__ perm, const int* __restrict__ totals, float* __restrict__ abuf) {
  using L = Layout<D>;
  const int nvalid = totals[kTotValid];
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < nvalid; j += gridDim.x * kThreads) {
    const int p = perm[j];
    const Params& P = scenes[pi[(int64_t)L::SCENE * cap + p]];
    const int mat = pi[(int64_t)L::MAT * cap + p];
    float C[3][3], F[3][3];
    load_cf<D>(pf, cap, p, C, F);
    Recompute<D> r;
    recompute<D>(P, mat, C, F, pf[(int64_t)L::JP * cap + p], P.youngs ? P.youngs[p] : 0.0f, r);
    const float iso = r.la * r.J * (r.J - 1.0f);
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) s += (r.Fp[a][k] - r.R[a][k]) * r.Fp[e][k];
        const float tau = 2.0f * r.mu * s + (a == e ? iso : 0.0f);
        abuf[(int64_t)(a * D + e) * cap + p] = P.stress_scale * tau + r.mass * C[a][e];
      }
  }
}

// Step 2 (per chunk, stencil): the g2p adjoint gather and the p2g adjoint
// gathers.  Writes xbar and vbar into gin, and Abar = dx sum_i w_i pbar_i
// d_i^T into gin's C rows (step 3 turns it into Cbar).
template <int D>
__global__ __launch_bounds__(kThreads) void k_p2g_adj(
    const Params* __restrict__ scenes, const int* __restrict__ slab,
    const float* __restrict__ pf, const int* __restrict__ pi, int64_t cap,
    const int* __restrict__ perm, const int4* __restrict__ chunks,
    const int* __restrict__ totals, const float4* __restrict__ grid,
    const float4* __restrict__ gbar, const float* __restrict__ gf,
    const float* __restrict__ abuf, float* __restrict__ gin, float* __restrict__ gE, int n) {
  using L = Layout<D>;
  constexpr int NPB = Geo<D>::kNodesPerBlock, TN = Geo<D>::kTileNodes;
  __shared__ float4 vt[TN];
  __shared__ float4 gt[TN];
  const int nchunks = totals[kTotChunks];
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int4 ck = chunks[c];
    int lb[3], lbl[3];
    decode_local<D>(ck.x, scenes[0], lb);
    const Params& P = scenes[scene_of<D>(lb, slab)];
    to_local<D>(lb, P, lbl);
    int b[3];
#pragma unroll
    for (int a = 0; a < D; a++) b[a] = (lbl[a] + P.blo[a]) * kBlock;
    // grid velocity after the grid operator, recomputed from the raw sums,
    // and the grid adjoint
    for (int i = threadIdx.x; i < TN; i += kThreads) {
      const int tl[3] = {i % kTile, (i / kTile) % kTile, i / (kTile * kTile)};
      int q[3], ln[3], I[3] = {0, 0, 0};
#pragma unroll
      for (int a = 0; a < D; a++) {
        const int hi = tl[a] >= kBlock;
        q[a] = lb[a] + hi;
        ln[a] = tl[a] - kBlock * hi;
        I[a] = b[a] + tl[a];
      }
      const int gb = encode_local<D>(q, P);
      const int node = D == 3 ? ln[0] + 4 * (ln[1] + 4 * ln[2]) : ln[0] + 4 * ln[1];
      const float4 f = grid[(int64_t)gb * NPB + node];
      float v[3] = {f.x, f.y, f.z};
      if (node_has_mass(f.w)) {
        const float inv = 1.0f / f.w;
#pragma unroll
        for (int a = 0; a < D; a++) v[a] = v[a] * inv + P.dt * P.gravity[a];
      }
      apply_post<D>(P, I, v);
      vt[i] = make_float4(v[0], v[1], D == 3 ? v[2] : 0.0f, 0.0f);
      gt[i] = gbar[(int64_t)gb * NPB + node];
    }
    __syncthreads();
    const int j = ck.y + threadIdx.x;
    if (j < ck.z) {
      const int p = perm[j];
      const int mat = pi[(int64_t)L::MAT * cap + p];
      float mass = P.p_mass;
      if (mat == kWater) mass *= P.water_density;
      float xp[3] = {0, 0, 0}, vp[3] = {0, 0, 0}, av[3] = {0, 0, 0}, xbo[3] = {0, 0, 0};
      float Cbo[3][3] = {{0}};
#pragma unroll
      for (int a = 0; a < D; a++) {
        xp[a] = pf[(int64_t)(L::X + a) * cap + p];
        vp[a] = pf[(int64_t)(L::V + a) * cap + p];
        xbo[a] = gf[(int64_t)(L::X + a) * cap + p];
        av[a] = gf[(int64_t)(L::V + a) * cap + p] + P.dt * xbo[a];
#pragma unroll
        for (int e = 0; e < D; e++) Cbo[a][e] = gf[(int64_t)(L::C + a * D + e) * cap + p];
      }
      float fx[3] = {0, 0, 0}, w[3][3], dw[3][3];
      int l[3] = {0, 0, 0};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        w[a][0] = w[a][2] = 0.0f;
        w[a][1] = 1.0f;
        dw[a][0] = dw[a][1] = dw[a][2] = 0.0f;
      }
#pragma unroll
      for (int a = 0; a < D; a++) {
        int base;
        base_fx(xp[a], P.inv_dx, base, fx[a]);
        bspline(fx[a], w[a]);
        bspline_grad(fx[a], dw[a]);
        l[a] = base - b[a];
      }
      float A[3][3] = {{0}};
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) A[a][e] = abuf[(int64_t)(a * D + e) * cap + p];
      float vnew[3] = {0, 0, 0}, xg[3] = {0, 0, 0}, S[3] = {0, 0, 0};
      float Ab[3][3] = {{0}};
#pragma unroll 1  // keeps the LDS loads of one z-layer in flight, not all 27
      for (int oz = 0; oz < (D == 3 ? 3 : 1); oz++)
#pragma unroll
        for (int oy = 0; oy < 3; oy++)
#pragma unroll
          for (int ox = 0; ox < 3; ox++) {
            const int o[3] = {ox, oy, oz};
            float wt = 1.0f, gw[3];
#pragma unroll
            for (int a = 0; a < D; a++) wt *= w[a][o[a]];
#pragma unroll
            for (int e = 0; e < D; e++) {
              float g = P.inv_dx * dw[e][o[e]];
#pragma unroll
              for (int a = 0; a < D; a++)
                if (a != e) g *= w[a][o[a]];
              gw[e] = g;
            }
            float d[3] = {0, 0, 0};
#pragma unroll
            for (int a = 0; a < D; a++) d[a] = (float)o[a] - fx[a];
            const int node = D == 3 ? (l[0] + ox) + kTile * ((l[1] + oy) + kTile * (l[2] + oz))
                                    : (l[0] + ox) + kTile * (l[1] + oy);
            const float4 v4 = vt[node], g4 = gt[node];
            const float vi[3] = {v4.x, v4.y, v4.z}, pb[3] = {g4.x, g4.y, g4.z};
            // s1: g2p adjoint (a . v_i + c4 v_i^T C'bar d_i);
            // s2: p2g adjoint pbar_i . (m v + A d_i dx) + m mbar_i
            float s1 = 0.0f, s2 = mass * g4.w;
#pragma unroll
            for (int a = 0; a < D; a++) {
              vnew[a] += wt * vi[a];
              float cd = 0.0f, ad = 0.0f;
#pragma unroll
              for (int e = 0; e < D; e++) {
                cd += Cbo[a][e] * d[e];
                ad += A[a][e] * d[e];
              }
              s1 += vi[a] * (av[a] + P.c4 * cd);
              s2 += pb[a] * (mass * vp[a] + P.dx * ad);
              S[a] += wt * pb[a];
#pragma unroll
              for (int e = 0; e < D; e++) Ab[a][e] += wt * pb[a] * d[e];
            }
#pragma unroll
            for (int e = 0; e < D; e++) xg[e] += (s1 + s2) * gw[e];
          }
#pragma unroll
      for (int e = 0; e < D; e++) {
        float t1 = 0.0f, t2 = 0.0f;
#pragma unroll
        for (int a = 0; a < D; a++) {
          t1 += Cbo[a][e] * vnew[a];
          t2 += A[a][e] * S[a];
        }
        gin[(int64_t)(L::X + e) * cap + p] = xbo[e] + xg[e] - P.c4 * P.inv_dx * t1 - t2;
        gin[(int64_t)(L::V + e) * cap + p] = mass * S[e];
      }
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) gin[(int64_t)(L::C + a * D + e) * cap + p] = P.dx * Ab[a][e];
    }
    __syncthreads();
  }
  // particles outside every chunk (escaped the window, sorted to the end)
  // took no part in the substep: zero adjoints
  for (int j = totals[kTotValid] + blockIdx.x * kThreads + threadIdx.x; j < n;
       j += gridDim.x * kThreads) {
    const int p = perm[j];
    for (int c = 0; c < L::NF; c++) gin[(int64_t)c * cap + p] = 0.0f;
    if (gE) gE[p] = 0.0f;
  }
}

// Step 3 (per particle): the constitutive adjoint.  Reads Abar from gin's C
// rows and writes Cbar, Fbar, Jpbar, Ebar and per-workgroup nu partials.
template <int D>
__global__ __launch_bounds__(kThreads) void k_adj_constitutive(
    const Params* __restrict__ scenes, const int* __restrict__ slab,
    const float* __restrict__ pf, const int* __restrict__ pi, int64_t cap,
    const int* __restrict__ perm, const int* __restrict__ totals, const float* __restrict__ gf,
    float* __restrict__ gin, float* __restrict__ gE, float* __restrict__ nu_part) {
  using L = Layout<D>;
  __shared__ float red[kThreads / 64];
  float nu_acc = 0.0f;
  const int nvalid = totals[kTotValid];
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < nvalid; j += gridDim.x * kThreads) {
    const int p = perm[j];
    const Params& P = scenes[pi[(int64_t)L::SCENE * cap + p]];
    const int mat = pi[(int64_t)L::MAT * cap + p];
    float C[3][3], F[3][3];
    load_cf<D>(pf, cap, p, C, F);
    const float Jp = pf[(int64_t)L::JP * cap + p];
    const float E = P.youngs ? P.youngs[p] : 0.0f;
    Recompute<D> r;
    recompute<D>(P, mat, C, F, Jp, E, r);
    float Cbar[3][3] = {{0}}, G[3][3] = {{0}}, Fbo[3][3] = {{0}};
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) {
        const float abar = gin[(int64_t)(L::C + a * D + e) * cap + p];
        Cbar[a][e] = r.mass * abar;
        G[a][e] = P.stress_scale * abar;
        Fbo[a][e] = gf[(int64_t)(L::F + a * D + e) * cap + p];
      }
    const float Jpbo = gf[(int64_t)L::JP * cap + p];
    float trG = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) trG += G[a][a];
    const float J = r.J;
    const float labar = trG * J * (J - 1.0f);
    float Jbar = r.la * trG * (2.0f * J - 1.0f);
    float Fbar[3][3] = {{0}}, Jpbar, mubar = 0.0f;
    if (r.water) {
      Jbar += Fbo[0][0] + Jpbo;
      const float detK = det_d<D>(r.K);
      float cK[3][3];
      cofactor_d<D>(r.K, cK);
      Jpbar = Jbar * detK;
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) Cbar[a][e] += Jbar * Jp * P.dt * cK[a][e];
    } else {
      const float mu = r.mu;
      float Fpb[3][3], Rb[3][3];
      {
        float cF[3][3];
        cofactor_d<D>(r.Fp, cF);
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
          for (int e = 0; e < D; e++) {
            float gfp = 0.0f, gtt = 0.0f, tft = 0.0f;
#pragma unroll
            for (int k = 0; k < D; k++) {
              gfp += G[a][k] * r.Fp[k][e];                          // (G F')_ae
              gtt += G[k][a] * (r.Fp[k][e] - r.R[k][e]);            // (G^T (F' - R))_ae
              tft += (r.Fp[a][k] - r.R[a][k]) * r.Fp[e][k];         // ((F' - R) F'^T)_ae
            }
            mubar += 2.0f * G[a][e] * tft;
            Fpb[a][e] = Fbo[a][e] + 2.0f * mu * (gfp + gtt) + Jbar * cF[a][e];
            Rb[a][e] = -2.0f * mu * gfp;
          }
      }
      // polar adjoint: B = U^T Rb V, Mbar = skew(B) / (sig_i + sig_j), F'bar += U Mbar V^T
      float Bm[3][3];
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          float t[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int m2 = 0; m2 < D; m2++)
#pragma unroll
            for (int k = 0; k < D; k++) t[m2] += r.U[k][a] * Rb[k][m2];
          float s = 0.0f;
#pragma unroll
          for (int m2 = 0; m2 < D; m2++) s += t[m2] * r.V[m2][e];
          Bm[a][e] = s;
        }
      float Mb[3][3];
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          if (a == e) {
            Mb[a][e] = 0.0f;
          } else {
            float den = r.sig[a] + r.sig[e];
            if (fabsf(den) < 1e-6f) den = den < 0.0f ? -1e-6f : 1e-6f;
            Mb[a][e] = (Bm[a][e] - Bm[e][a]) / den;
          }
        }
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          float s = 0.0f;
#pragma unroll
          for (int k = 0; k < D; k++)
#pragma unroll
            for (int m2 = 0; m2 < D; m2++) s += r.U[a][k] * Mb[k][m2] * r.V[e][m2];
          Fpb[a][e] += s;
        }
      // F' = K F: Fbar = K^T F'bar, Cbar += dt F'bar F^T
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          float s = 0.0f, t = 0.0f;
#pragma unroll
          for (int k = 0; k < D; k++) {
            s += r.K[k][a] * Fpb[k][e];
            t += Fpb[a][k] * F[e][k];
          }
          Fbar[a][e] = s;
          Cbar[a][e] += P.dt * t;
        }
      Jpbar = Jpbo;
    }
    // mu = h E kmu, la = h E kla
    if (gE) gE[p] = r.h * (P.kmu * mubar + P.kla * labar);
    nu_acc += r.h * E * (P.dkmu * mubar + P.dkla * labar);
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int e = 0; e < D; e++) {
        gin[(int64_t)(L::C + a * D + e) * cap + p] = Cbar[a][e];
        gin[(int64_t)(L::F + a * D + e) * cap + p] = Fbar[a][e];
      }
    gin[(int64_t)L::JP * cap + p] = Jpbar;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nu_acc += __shfl_xor(nu_acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nu_acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) s += red[w];
    nu_part[blockIdx.x] = s;
  }
}

// Sum of the per-workgroup nu partials in a fixed order (one workgroup).
__global__ __launch_bounds__(kThreads) void k_sum_partials(const float* __restrict__ part, int n,
                                                           float* __restrict__ out) {
  __shared__ float red[kThreads / 64];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += kThreads) s += part[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) t += red[w];
    out[0] = t;
  }
}

}  // namespace mpm
