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
//
// Snow and sand (plastic_adjoint below).  Both store F = U diag(sh) V^T
// with sh a function of the singular values sig of F' = K F (snow: clamped,
// mpm:527-545; sand: the return map, mpm:321-342, 552-556), and both
// stresses are tau = U diag(t) U^T (snow: t_i = 2 mu (sh_i - 1) sh_i +
// la J (J - 1), J = prod sh; sand: t_i = 2 mu_0 l_i + lambda_0 sum l,
// l = ln sh).  With M = U^T Fbar_out V and N = sym(U^T taubar U), the
// adjoint of F' is U X V^T where, for i != j,
//   X_ij = sym(M)_ij Dg_ij + skew(M)_ij (sh_i + sh_j) / (sig_i + sig_j)
//          + 2 N_ij sig_j Dt_ij / (sig_i + sig_j),
// Dg_ij = (sh_j - sh_i) / (sig_j - sig_i) and Dt_ij = (t_j - t_i) /
// (sig_j - sig_i) being divided differences taken in closed form (no
// 1 / (sig_i^2 - sig_j^2) of the SVD's own derivative, so repeated singular
// values, F' = I included, are harmless), and X_ii = sigbar_i, the adjoint
// of the singular values through sh, t and Jp.  These were checked against
// float64 finite differences of the forward in every branch (snow inside /
// at its clamp; sand at the cone tip, inside the cone, projected).
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

// expm1(a d) / expm1(d), continuous through d = 0 (value a)
__device__ __forceinline__ float ratio_expm1(float a, float d) {
  if (fabsf(d) < 1e-4f) return a * (1.0f + 0.5f * (a - 1.0f) * d);
  return expm1f(a * d) / expm1f(d);
}

// d / expm1(d), continuous through d = 0 (value 1)
__device__ __forceinline__ float ratio_lin_expm1(float d) {
  if (fabsf(d) < 1e-4f) return 1.0f - 0.5f * d;
  return d / expm1f(d);
}

// Constitutive adjoint of one snow or sand particle (see the header): given
// F' = K F, Jp, E and the adjoints G = taubar, Fbo = Fbar_out, Jpbo =
// Jpbar_out, returns Fpbar (the adjoint of F'), Jpbar, and the adjoints of
// mu_0 and lambda_0 (and, for sand, of the return map's coefficient
// c = sand_coef alpha, through cbar).
template <int D>
__device__ __forceinline__ void plastic_adjoint(const Params& P, int mat, const float Fp[3][3],
                                                float Jp, float mu0, float la0,
                                                const float G[3][3], const float Fbo[3][3],
                                                float Jpbo, float Fpb[3][3], float& Jpbar,
                                                float& mu0b, float& la0b, float& cbar) {
  float U[3][3], V[3][3], sig[3] = {1.0f, 1.0f, 1.0f};
  if (D == 3)
    svd3(Fp, U, sig, V);
  else
    svd2(Fp, U, sig, V);
  // N = sym(U^T G U), M = U^T Fbo V
  float N[3][3] = {{0.0f}}, M[3][3] = {{0.0f}};
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      float n = 0.0f, m = 0.0f;
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          n += U[a][i] * G[a][e] * U[e][j];
          m += U[a][i] * Fbo[a][e] * V[e][j];
        }
      N[i][j] = n;
      M[i][j] = m;
    }
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < i; j++) N[i][j] = N[j][i] = 0.5f * (N[i][j] + N[j][i]);
  float sh[3] = {1.0f, 1.0f, 1.0f}, sb[3] = {0.0f, 0.0f, 0.0f};
  float Dg[3][3] = {{0.0f}}, Dt[3][3] = {{0.0f}};
  float trN = 0.0f;
#pragma unroll
  for (int i = 0; i < D; i++) trN += N[i][i];
  cbar = 0.0f;
  if (mat == kSnow) {
    const float lo = 1.0f - 2.5e-2f, hi = 1.0f + 4.5e-3f;  // mpm_solver.py:530-531
    const float h = __expf(10.0f * (1.0f - Jp));            // mpm_solver.py:515-518
    const float mu = mu0 * h, la = la0 * h;
    float gp[3] = {0.0f, 0.0f, 0.0f}, J = 1.0f, r = 1.0f;
#pragma unroll
    for (int i = 0; i < D; i++) {
      gp[i] = (sig[i] > lo && sig[i] < hi) ? 1.0f : 0.0f;
      sh[i] = fminf(fmaxf(sig[i], lo), hi);
      J *= sh[i];
      r *= sig[i] / sh[i];
    }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        if (i == j) continue;
        float dg;
        if (gp[i] != 0.0f && gp[j] != 0.0f)
          dg = 1.0f;
        else if (fabsf(sig[j] - sig[i]) > 1e-6f)
          dg = (sh[j] - sh[i]) / (sig[j] - sig[i]);
        else
          dg = 0.5f * (gp[i] + gp[j]);
        Dg[i][j] = dg;
        Dt[i][j] = 2.0f * mu * (sh[i] + sh[j] - 1.0f) * dg;
      }
    const float Jpn = Jp * r;
    float mub = 0.0f;
#pragma unroll
    for (int i = 0; i < D; i++) {
      float jo = 1.0f;
#pragma unroll
      for (int k = 0; k < D; k++)
        if (k != i) jo *= sh[k];
      const float shb = M[i][i] + 2.0f * mu * (2.0f * sh[i] - 1.0f) * N[i][i] +
                        la * (2.0f * J - 1.0f) * jo * trN;
      sb[i] = gp[i] * shb + Jpbo * Jpn * (1.0f / sig[i] - gp[i] / sh[i]);
      mub += N[i][i] * 2.0f * (sh[i] - 1.0f) * sh[i];
    }
    const float lab = trN * J * (J - 1.0f);
    const float hb = mu0 * mub + la0 * lab;
    Jpbar = Jpbo * r - 10.0f * h * hb;
    mu0b = h * mub;
    la0b = h * lab;
  } else {  // sand
    const float c = P.sand_coef * P.alpha;
    float eps[3] = {0.0f, 0.0f, 0.0f}, epsd[3] = {0.0f, 0.0f, 0.0f}, tr = Jp;
#pragma unroll
    for (int i = 0; i < D; i++) {
      eps[i] = logf(fmaxf(fabsf(sig[i]), 1e-4f));
      epsd[i] = fabsf(sig[i]) > 1e-4f ? 1.0f / sig[i] : 0.0f;
      tr += eps[i];
    }
    float hat[3] = {0.0f, 0.0f, 0.0f}, n2 = 0.0f;
#pragma unroll
    for (int i = 0; i < D; i++) {
      hat[i] = eps[i] - tr / (float)D;
      n2 += hat[i] * hat[i];
    }
    const float n = sqrtf(n2) + 1e-20f;
    // l = a eps + (common shift): tip a = 0 (sh = 1), inside the cone
    // a = 1 (sh = |sig|), projected a = 1 - k
    int branch = 0;  // 0 tip, 1 inside, 2 projected
    float a = 0.0f, ell[3] = {0.0f, 0.0f, 0.0f};
    if (tr < 0.0f) {
      const float dg = n + c * tr;
      if (dg <= 0.0f) {
        branch = 1;
        a = 1.0f;
#pragma unroll
        for (int i = 0; i < D; i++) ell[i] = eps[i];
      } else {
        branch = 2;
        const float k = dg / n;
        a = 1.0f - k;
#pragma unroll
        for (int i = 0; i < D; i++) ell[i] = eps[i] - k * hat[i];
      }
#pragma unroll
      for (int i = 0; i < D; i++) sh[i] = expf(ell[i]);
    }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        if (i == j) continue;
        const float de = eps[j] - eps[i];
        Dg[i][j] = sh[i] / sig[i] * ratio_expm1(a, de);
        Dt[i][j] = 2.0f * mu0 * a / sig[i] * ratio_lin_expm1(de);
      }
    float ellb[3] = {0.0f, 0.0f, 0.0f}, Sl = 0.0f, lsum = 0.0f;
    mu0b = 0.0f;
#pragma unroll
    for (int i = 0; i < D; i++) {
      ellb[i] = M[i][i] * sh[i] + 2.0f * mu0 * N[i][i] + la0 * trN;
      Sl += ellb[i];
      lsum += ell[i];
      mu0b += N[i][i] * 2.0f * ell[i];
    }
    la0b = trN * lsum;
    float epsb[3] = {0.0f, 0.0f, 0.0f};
    if (branch == 0) {  // Jp' = tr, sh = 1
#pragma unroll
      for (int i = 0; i < D; i++) epsb[i] = Jpbo;
      Jpbar = Jpbo;
    } else if (branch == 1) {  // Jp' = 0, l = eps
#pragma unroll
      for (int i = 0; i < D; i++) epsb[i] = ellb[i];
      Jpbar = 0.0f;
    } else {  // Jp' = 0, l_i = tr / d - c tr u_i with u = hat / |hat|
      float u[3] = {0.0f, 0.0f, 0.0f}, lu = 0.0f, su = 0.0f;
#pragma unroll
      for (int i = 0; i < D; i++) {
        u[i] = hat[i] / n;
        lu += ellb[i] * u[i];
        su += u[i];
      }
      const float ctn = c * tr / n, invd = 1.0f / (float)D;
#pragma unroll
      for (int i = 0; i < D; i++)
        epsb[i] = Sl * invd - c * lu - ctn * (ellb[i] - Sl * invd - lu * (u[i] - su * invd));
      Jpbar = Sl * invd - c * lu + ctn * invd * (Sl - lu * su);
      cbar = -tr * lu;
    }
#pragma unroll
    for (int i = 0; i < D; i++) sb[i] = epsb[i] * epsd[i];
  }
  // X in the (U, V) frame, then Fpbar = U X V^T
  float X[3][3] = {{0.0f}};
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      if (i == j) {
        X[i][i] = sb[i];
        continue;
      }
      float den = sig[i] + sig[j];
      if (fabsf(den) < 1e-6f) den = den < 0.0f ? -1e-6f : 1e-6f;
      X[i][j] = 0.5f * (M[i][j] + M[j][i]) * Dg[i][j] +
                (0.5f * (M[i][j] - M[j][i]) * (sh[i] + sh[j]) + 2.0f * N[i][j] * sig[j] * Dt[i][j]) /
                    den;
    }
#pragma unroll
  for (int a = 0; a < D; a++)
#pragma unroll
    for (int e = 0; e < D; e++) {
      float s = 0.0f;
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) s += U[a][i] * X[i][j] * V[e][j];
      Fpb[a][e] = s;
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
    if (mat == kSnow || mat == kSand) {  // the forward's own constitutive update
      float Jc = pf[(int64_t)L::JP * cap + p], stress[3][3];
      const float E = P.youngs ? P.youngs[p] : 0.0f;
      constitutive<D>(mat, P, P.youngs ? E * P.kmu : P.mu_0, P.youngs ? E * P.kla : P.lambda_0, F,
                      C, Jc, stress);
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++)
          abuf[(int64_t)(a * D + e) * cap + p] = P.stress_scale * stress[a][e] + P.p_mass * C[a][e];
      continue;
    }
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
    for (int j = ck.y + threadIdx.x; j < ck.z; j += kThreads) {  // the chunk, kThreads at a time
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
    if (mat == kSnow || mat == kSand) {
      float K[3][3] = {{0.0f}}, Fp[3][3] = {{0.0f}}, G[3][3] = {{0.0f}}, Fbo[3][3] = {{0.0f}};
      float Cb[3][3] = {{0.0f}};
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          K[a][e] = (a == e ? 1.0f : 0.0f) + P.dt * C[a][e];
          const float abar = gin[(int64_t)(L::C + a * D + e) * cap + p];
          Cb[a][e] = P.p_mass * abar;
          G[a][e] = P.stress_scale * abar;
          Fbo[a][e] = gf[(int64_t)(L::F + a * D + e) * cap + p];
        }
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          float s = 0.0f;
#pragma unroll
          for (int k = 0; k < D; k++) s += K[a][k] * F[k][e];
          Fp[a][e] = s;
        }
      float Fpb[3][3], Jpbar, mu0b, la0b, cb;
      plastic_adjoint<D>(P, mat, Fp, Jp, P.youngs ? E * P.kmu : P.mu_0,
                         P.youngs ? E * P.kla : P.lambda_0, G, Fbo,
                         gf[(int64_t)L::JP * cap + p], Fpb, Jpbar, mu0b, la0b, cb);
      // F' = K F: Fbar = K^T F'bar, Cbar += dt F'bar F^T
#pragma unroll
      for (int a = 0; a < D; a++)
#pragma unroll
        for (int e = 0; e < D; e++) {
          float fb = 0.0f, t = 0.0f;
#pragma unroll
          for (int k = 0; k < D; k++) {
            fb += K[k][a] * Fpb[k][e];
            t += Fpb[a][k] * F[e][k];
          }
          gin[(int64_t)(L::C + a * D + e) * cap + p] = Cb[a][e] + P.dt * t;
          gin[(int64_t)(L::F + a * D + e) * cap + p] = fb;
        }
      gin[(int64_t)L::JP * cap + p] = Jpbar;
      // mu_0 = E kmu, lambda_0 = E kla; sand's coefficient depends on nu
      if (gE) gE[p] = P.kmu * mu0b + P.kla * la0b;
      nu_acc += E * (P.dkmu * mu0b + P.dkla * la0b) + cb * P.dsand_coef * P.alpha;
      continue;
    }
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
