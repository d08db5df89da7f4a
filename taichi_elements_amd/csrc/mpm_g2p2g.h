// use_g2p2g mode: the fused transfer of mpm_solver.py:363-485, device code
// included by mpm_kernels.hip.
//
// One work item is a chunk of the particles binned (by their current x) in
// block b.  The workgroup gathers the input grid's (4+2)^d node tile around
// b (G2P: v', C'; particles added since the last substep skip it and keep v
// with C' = 0), advects the particles (not the stationary ones), applies the
// constitutive update with C' (the particle's C field is not written, as in
// the reference), and scatters the new momenta around the NEW positions into
// an (4+4)^d node tile whose origin is one node below the block: a particle
// moves less than one cell per substep under the grid velocity clamp, so its
// new stencil stays inside.  Particles whose new cell lies in the block's own
// 4^d cells go through the cell-major conflict-free LDS scatter of p2g; the
// few that crossed into a neighbour's cells are added with LDS float atomics
// afterwards.  The tile goes to HBM, and the grid operator (MODE 2) sums the
// tiles of the 3^d neighbouring blocks into the output grid.  The same pass
// writes the next substep's binning hints (direction codes and ranks).
#pragma once

#include "mpm_device.h"

namespace mpm {

template <int D>
struct G2P2GGeo {
  static constexpr int kEdge = kBlock + 4;  // output tile nodes per axis
  static constexpr int kNodes = D == 3 ? kEdge * kEdge * kEdge : kEdge * kEdge;
};

template <int D, bool Q>
__global__ __launch_bounds__(kThreads) void k_g2p2g(
    const Params* __restrict__ scenes, const int* __restrict__ slab, float* __restrict__ pf,
    const int* __restrict__ pi, int64_t cap, const int* __restrict__ perm,
    const int4* __restrict__ chunks, const int* __restrict__ totals,
    const float4* __restrict__ grid_in, float4* __restrict__ tiles, int n_final,
    unsigned* __restrict__ maxv, const int* __restrict__ step, int stride,
    int* __restrict__ code, int* __restrict__ nkey,
    int* __restrict__ dcount, int* __restrict__ count_next, int* __restrict__ err) {
  using L = Layout<D>;
  constexpr int NPB = Geo<D>::kNodesPerBlock, TN = Geo<D>::kTileNodes;
  constexpr int E8 = G2P2GGeo<D>::kEdge, TN8 = G2P2GGeo<D>::kNodes;
  constexpr int CELLS = NPB;
  constexpr int GROUPS = 64 / CELLS;
  constexpr int WAVES = kThreads / 64;
  constexpr int SUBS = WAVES * GROUPS;
  __shared__ float tin_x[TN], tin_y[TN], tin_z[TN];  // input grid velocity (as k_g2p)
  __shared__ float4 stage[4][kBatch];  // [record part][particle]: lane-contiguous
  __shared__ float4 ptile[SUBS][TN];  // inner particles: nodes 1..6 of the output tile
  __shared__ float4 stile[TN8];       // particles that crossed a cell face (LDS atomics)
  __shared__ int wcnt[WAVES][CELLS];
  __shared__ int cstart[CELLS], ccount[CELLS];
  __shared__ unsigned char order[kBatch];
  __shared__ int dwcnt[WAVES][32];
  __shared__ float red[WAVES];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  float vmax = 0.0f;
  const int nchunks = totals[kTotChunks];
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int4 ck = chunks[c];
    const int nc = ck.z - ck.y;
    int lb[3], lbl[3], b[3] = {0, 0, 0};
    decode_local<D>(ck.x, scenes[0], lb);
    const Params& P = scenes[scene_of<D>(lb, slab)];
    to_local<D>(lb, P, lbl);
#pragma unroll
    for (int a = 0; a < D; a++) b[a] = (lbl[a] + P.blo[a]) * kBlock;
    for (int i = threadIdx.x; i < SUBS * TN; i += kThreads)
      (&ptile[0][0])[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int i = threadIdx.x; i < TN8; i += kThreads) stile[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int i = threadIdx.x; i < WAVES * CELLS; i += kThreads) (&wcnt[0][0])[i] = 0;
    for (int i = threadIdx.x; i < WAVES * 32; i += kThreads) (&dwcnt[0][0])[i] = 0;
    for (int i = threadIdx.x; i < TN; i += kThreads) {
      const int tl[3] = {i % kTile, (i / kTile) % kTile, i / (kTile * kTile)};
      int q[3], ln[3];
#pragma unroll
      for (int a = 0; a < D; a++) {
        const int hi = tl[a] >= kBlock;
        q[a] = lb[a] + hi;
        ln[a] = tl[a] - kBlock * hi;
      }
      const int gb = encode_local<D>(q, P);
      const int node = D == 3 ? ln[0] + 4 * (ln[1] + 4 * ln[2]) : ln[0] + 4 * ln[1];
      const float4 g = grid_in[(int64_t)gb * NPB + node];
      tin_x[i] = g.x;
      tin_y[i] = g.y;
      tin_z[i] = g.z;
    }
    __syncthreads();

    // ---- phase 1: G2P, advection, constitutive update, staging ----
    const int t = threadIdx.x;
    const bool live = t < nc;
    bool inner = false;
    int cell = 0, shell = -1;  // shell: output-tile node of the new base (crossed particles)
    int dir = 13, key_new = ck.x;
    if (live) {
      using PS = Particles<D, Q, true>;
      const int p = perm[ck.y + t];
      const int mat = pi[(int64_t)L::MAT * cap + p];
      float xp[3] = {0, 0, 0}, fx[3] = {0, 0, 0}, w[3][3], vp[3] = {0, 0, 0};
      int l[3] = {0, 0, 0};
#pragma unroll
      for (int a = 0; a < 3; a++) w[a][0] = w[a][1] = w[a][2] = 0.0f;
      PS::load_x(pf, cap, p, xp);
      PS::load_v(pf, cap, p, vp);
#pragma unroll
      for (int a = 0; a < D; a++) {
        int base;
        base_fx(xp[a], P.inv_dx, base, fx[a]);
        bspline(fx[a], w[a]);
        l[a] = min(max(base - b[a], 0), kBlock - 1);
      }
      float nv[3] = {0.0f, 0.0f, 0.0f}, nC[3][3] = {{0.0f}};
      const bool fresh = pi[(int64_t)L::ORIG * cap + p] >= n_final;  // mpm:396-399
      if (!fresh) {  // G2P, mpm:386-394 (tensor-product form of k_g2p)
        float pe[3][3] = {{0.0f}}, wd[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int k = 0; k < 3; k++) wd[a][k] = a < D ? ((float)k - fx[a]) * w[a][k] : 0.0f;
#pragma unroll
        for (int oz = 0; oz < (D == 3 ? 3 : 1); oz++) {
          float b0[3] = {0, 0, 0}, bx[3] = {0, 0, 0}, by[3] = {0, 0, 0};
#pragma unroll
          for (int oy = 0; oy < 3; oy++) {
            float a0[3] = {0, 0, 0}, ax[3] = {0, 0, 0};
#pragma unroll
            for (int ox = 0; ox < 3; ox++) {
              const int node = D == 3 ? (l[0] + ox) + kTile * ((l[1] + oy) + kTile * (l[2] + oz))
                                      : (l[0] + ox) + kTile * (l[1] + oy);
              const float gg[3] = {tin_x[node], tin_y[node], D == 3 ? tin_z[node] : 0.0f};
#pragma unroll
              for (int cc = 0; cc < D; cc++) {
                a0[cc] += w[0][ox] * gg[cc];
                ax[cc] += wd[0][ox] * gg[cc];
              }
            }
#pragma unroll
            for (int cc = 0; cc < D; cc++) {
              b0[cc] += w[1][oy] * a0[cc];
              bx[cc] += w[1][oy] * ax[cc];
              by[cc] += wd[1][oy] * a0[cc];
            }
          }
          const float wz = D == 3 ? w[2][oz] : 1.0f;
#pragma unroll
          for (int cc = 0; cc < D; cc++) {
            nv[cc] += wz * b0[cc];
            pe[cc][0] += wz * bx[cc];
            pe[cc][1] += wz * by[cc];
            if (D == 3) pe[cc][2] += wd[2][oz] * b0[cc];
          }
        }
#pragma unroll
        for (int cc = 0; cc < D; cc++)
#pragma unroll
          for (int e = 0; e < D; e++) nC[cc][e] = P.c4 * pe[cc][e];  // no cancellation (see k_g2p)
      } else {
#pragma unroll
        for (int a = 0; a < D; a++) nv[a] = vp[a];
      }
      float xn[3] = {xp[0], xp[1], xp[2]};
      if (mat != kStationary) {  // mpm:401-403: x advances by the stored v
        PS::store_v(pf, cap, p, nv);
#pragma unroll
        for (int a = 0; a < D; a++) {
          vp[a] = nv[a];
          xn[a] = xp[a] + P.dt * nv[a];
        }
        PS::store_x(pf, cap, p, xn);  // xn: the stored x, scattered around below
      }
#pragma unroll
      for (int a = 0; a < D; a++) vmax = fmaxf(vmax, fabsf(vp[a]));
      // new base relative to the output tile origin (one node below the block)
      float fn[3] = {0, 0, 0};
      int rel[3] = {1, 1, 1};
      bool fits = true;
#pragma unroll
      for (int a = 0; a < D; a++) {
        int base;
        base_fx(xn[a], P.inv_dx, base, fn[a]);
        rel[a] = base - (b[a] - 1);
        fits &= (rel[a] >= 0) & (rel[a] <= E8 - 3);
      }
      if (!fits) {
        atomicOr(err, 8);  // moved more than one cell in a substep
      } else {
        inner = true;
#pragma unroll
        for (int a = 0; a < D; a++) inner &= (rel[a] >= 1) & (rel[a] <= kBlock);
        if (inner)
          cell = D == 3 ? (rel[0] - 1) + 4 * ((rel[1] - 1) + 4 * (rel[2] - 1))
                        : (rel[0] - 1) + 4 * (rel[1] - 1);
        else
          shell = D == 3 ? rel[0] + E8 * (rel[1] + E8 * rel[2]) : rel[0] + E8 * rel[1];
      }
      // constitutive update with the G2P's C (mpm:409-473)
      float F[3][3] = {{0}}, stress[3][3] = {{0}};
      PS::load_F(pf, cap, p, F);
      float Jp = P.plastic ? PS::load_Jp(pf, cap, p) : 1.0f;
      constitutive<D, true, Q>(mat, P, P.mu_0, P.lambda_0, F, nC, Jp, stress);
      PS::store_F(pf, cap, p, F);
      if (P.plastic) PS::store_Jp(pf, cap, p, Jp);
      const float mass = P.p_mass;  // mpm:474-483 (no water density here)
      float Ad[3][3] = {{0.0f}}, q[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int a = 0; a < D; a++) {
#pragma unroll
        for (int e = 0; e < D; e++) Ad[a][e] = (P.stress_scale * stress[a][e] + mass * nC[a][e]) * P.dx;
        float s = mass * vp[a];
#pragma unroll
        for (int e = 0; e < D; e++) s -= Ad[a][e] * fn[e];
        q[a] = s;
      }
      if (!fits) {  // contributes nothing; the caller is told
#pragma unroll
        for (int a = 0; a < 3; a++) {
          q[a] = 0.0f;
#pragma unroll
          for (int e = 0; e < 3; e++) Ad[a][e] = 0.0f;
        }
      }
      if (D == 3) {
        stage[0][t] = make_float4(Ad[0][0], Ad[0][1], Ad[0][2], Ad[1][0]);
        stage[1][t] = make_float4(Ad[1][1], Ad[1][2], Ad[2][0], Ad[2][1]);
        stage[2][t] = make_float4(Ad[2][2], q[0], q[1], q[2]);
        stage[3][t] = make_float4(fn[0], fn[1], fn[2], fits ? mass : 0.0f);
      } else {
        stage[0][t] = make_float4(Ad[0][0], Ad[0][1], Ad[1][0], Ad[1][1]);
        stage[1][t] = make_float4(q[0], q[1], fn[0], fn[1]);
        stage[2][t] = make_float4(fits ? mass : 0.0f, 0.0f, 0.0f, 0.0f);
      }
      // block of the new position for the next substep's binning
      int dd[3] = {0, 0, 0}, qb[3] = {0, 0, 0};
      bool near = true;
#pragma unroll
      for (int a = 0; a < D; a++) {
        qb[a] = ((b[a] - 1 + rel[a]) >> 2) - P.blo[a];
        dd[a] = qb[a] - lbl[a];
        near &= (dd[a] >= -1) & (dd[a] <= 1) & (qb[a] >= 0) & (qb[a] < P.bnl[a] - 1);
      }
      if (fits && near) {
        dir = (dd[2] + 1) * 9 + (dd[1] + 1) * 3 + (dd[0] + 1);
        qb[D - 1] += P.zoff;
        key_new = encode_local<D>(qb, P);
      } else if (fits) {
        atomicOr(err, 1);  // left the window
      }
    }
    __syncthreads();

    // ---- phase 2: stable counting sort of the inner particles by cell ----
    const bool in_cell = live && inner;
    unsigned long long peers = __ballot(in_cell);
#pragma unroll
    for (int bit = 0; bit < (D == 3 ? 6 : 4); bit++) {
      const bool on = (cell >> bit) & 1;
      const unsigned long long bal = __ballot(on);
      peers &= on ? bal : ~bal;
    }
    const int rank = (int)__popcll(peers & lt);
    if (in_cell && rank == 0) wcnt[wave][cell] = (int)__popcll(peers);
    __syncthreads();
    if (threadIdx.x < 64) {
      int run = 0;
      if (lane < CELLS) {
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
          const int cnt = wcnt[w][lane];
          wcnt[w][lane] = run;
          run += cnt;
        }
      }
      int inc = run;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
      }
      if (lane < CELLS) {
        ccount[lane] = run;
        cstart[lane] = inc - run;
      }
    }
    __syncthreads();
    if (in_cell) order[cstart[cell] + wcnt[wave][cell] + rank] = (unsigned char)t;
    __syncthreads();

    // ---- phase 3: cell-major scatter of the inner particles ----
    {
      const int my_cell = scatter_cell_of_lane<D>(lane);
      const int sub = wave * GROUPS + lane / CELLS;
      const int lc[3] = {my_cell & 3, (my_cell >> 2) & 3, my_cell >> 4};
      const int cbeg = cstart[my_cell], ccnt = ccount[my_cell];
      float4* tp = ptile[sub];
      cell_scatter<D, SUBS>(sub, ccnt, cbeg, order, stage, lc, tp);
    }
    __syncthreads();

    // ---- phase 3b: particles that crossed into a neighbour's cells ----
    if (live && shell >= 0) {
      float Ad[3][3], q[3], mass, w[3][3], f3[3];
      const float4 s0 = stage[0][t], s1 = stage[1][t], s2 = stage[2][t], s3 = stage[3][t];
      if (D == 3) {
        Ad[0][0] = s0.x; Ad[0][1] = s0.y; Ad[0][2] = s0.z;
        Ad[1][0] = s0.w; Ad[1][1] = s1.x; Ad[1][2] = s1.y;
        Ad[2][0] = s1.z; Ad[2][1] = s1.w; Ad[2][2] = s2.x;
        q[0] = s2.y; q[1] = s2.z; q[2] = s2.w;
        f3[0] = s3.x; f3[1] = s3.y; f3[2] = s3.z;
        mass = s3.w;
      } else {
        Ad[0][0] = s0.x; Ad[0][1] = s0.y; Ad[1][0] = s0.z; Ad[1][1] = s0.w;
        Ad[0][2] = Ad[1][2] = Ad[2][0] = Ad[2][1] = Ad[2][2] = 0.0f;
        q[0] = s1.x; q[1] = s1.y; q[2] = 0.0f;
        f3[0] = s1.z; f3[1] = s1.w; f3[2] = 0.0f;
        mass = s2.x;
      }
#pragma unroll
      for (int a = 0; a < 3; a++) bspline(f3[a], w[a]);
      float4* tp = stile;
#pragma unroll
      for (int oz = 0; oz < (D == 3 ? 3 : 1); oz++)
#pragma unroll
        for (int oy = 0; oy < 3; oy++)
#pragma unroll
          for (int ox = 0; ox < 3; ox++) {
            float weight = w[0][ox] * w[1][oy];
            if (D == 3) weight *= w[2][oz];
            const float fo[3] = {(float)ox, (float)oy, (float)oz};
            float m3[3];
#pragma unroll
            for (int a = 0; a < 3; a++)
              m3[a] = q[a] + fo[0] * Ad[a][0] + fo[1] * Ad[a][1] + fo[2] * Ad[a][2];
            const int node = shell + (D == 3 ? ox + E8 * (oy + E8 * oz) : ox + E8 * oy);
            atomicAdd(&tp[node].x, weight * m3[0]);
            atomicAdd(&tp[node].y, weight * m3[1]);
            if (D == 3) atomicAdd(&tp[node].z, weight * m3[2]);
            atomicAdd(&tp[node].w, weight * mass);
          }
    }

    // ---- next substep's binning hints (as k_g2p) ----
    unsigned long long dpeers = __ballot(live);
#pragma unroll
    for (int bit = 0; bit < 5; bit++) {
      const bool on = (dir >> bit) & 1;
      const unsigned long long bal = __ballot(on);
      dpeers &= on ? bal : ~bal;
    }
    const int drank = (int)__popcll(dpeers & lt);
    if (live && drank == 0) dwcnt[wave][dir] = (int)__popcll(dpeers);
    __syncthreads();
    if (threadIdx.x < 27) {
      const int d = threadIdx.x;
      int run = 0;
#pragma unroll
      for (int w = 0; w < WAVES; w++) {
        const int cnt = dwcnt[w][d];
        dwcnt[w][d] = run;
        run += cnt;
      }
      dcount[(int64_t)c * 27 + d] = run;
      if (run > 0) {
        int q[3] = {lb[0] + d % 3 - 1, lb[1] + (d / 3) % 3 - 1, lb[2] + d / 9 - 1};
        atomicAdd(&count_next[encode_local<D>(q, P)], run);
      }
    }

    // ---- phase 4: sum the private tiles, write the node tile ----
    float4* out = tiles + (int64_t)c * TN8;
    for (int i = threadIdx.x; i < TN8; i += kThreads) {
      const int t8[3] = {i % E8, (i / E8) % E8, i / (E8 * E8)};
      bool in6 = true;
#pragma unroll
      for (int a = 0; a < D; a++) in6 &= (t8[a] >= 1) & (t8[a] <= kTile);
      float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (in6) {
        const int j6 = D == 3 ? (t8[0] - 1) + kTile * ((t8[1] - 1) + kTile * (t8[2] - 1))
                              : (t8[0] - 1) + kTile * (t8[1] - 1);
#pragma unroll
        for (int sdx = 0; sdx < SUBS; sdx++) {
          const float4 v = ptile[sdx][j6];
          acc.x += v.x;
          acc.y += v.y;
          acc.z += v.z;
          acc.w += v.w;
        }
      }
      const float4 sv = stile[i];
      acc.x += sv.x;
      acc.y += sv.y;
      acc.z += sv.z;
      acc.w += sv.w;
      out[i] = acc;
    }
    __syncthreads();
    if (live) {
      code[ck.y + t] = dir | ((dwcnt[wave][dir] + drank) << 8);
      nkey[ck.y + t] = key_new;
    }
    __syncthreads();
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
#pragma unroll
    for (int w = 1; w < WAVES; w++) m = fmaxf(m, red[w]);
    maxv[(int64_t)(step ? *step : 0) * stride + blockIdx.x] = __float_as_uint(m);
  }
}

// Grid operator of the use_g2p2g mode: the node sums of the (4+4)^d tiles of
// the chunks of blocks b + {-1, 0, 1}^d, then normalisation and gravity
// (mpm:591-593), the velocity clamp (mpm:596-598) and the post-processes.
template <int D>
__global__ __launch_bounds__(kThreads) void k_grid_op_g2p2g(
    const Params* __restrict__ scenes, const int* __restrict__ slab,
    const int* __restrict__ active, const int* __restrict__ totals,
    const int* __restrict__ pchunk_begin, const int* __restrict__ pchunk_cnt,
    const float4* __restrict__ tiles, float4* __restrict__ grid) {
  constexpr int NPB = Geo<D>::kNodesPerBlock;
  constexpr int E8 = G2P2GGeo<D>::kEdge, TN8 = G2P2GGeo<D>::kNodes;
  // The tile of block q = b - d starts one node below q, so this node sits at
  // tile coordinate l + 1 + 4 d, inside the tile for d = 0 and, per axis,
  // for d = -1 (l = 3) or d = +1 (l < 3): 2^D source blocks, summed by
  // node_lanes lanes per node and combined in a fixed tree as in k_grid_op.
  constexpr int SRC = D == 3 ? 8 : 4;
  const int64_t nodes = (int64_t)totals[kTotActive] * NPB;
  const int lanes = node_lanes(nodes, SRC);
  for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < nodes * lanes;
       t += (int64_t)gridDim.x * kThreads) {
    const int64_t nt = t / lanes;
    const int sub = (int)(t % lanes);
    const int blk = (int)(nt / NPB), node = (int)(nt % NPB);
    const int bid = active[blk];
    int lb[3];
    decode_local<D>(bid, scenes[0], lb);
    const Params& P = scenes[scene_of<D>(lb, slab)];
    if (P.skip) continue;  // uniform over the group
    const int l[3] = {node & 3, (node >> 2) & 3, node >> 4};
    const float4 sum = tree_sum<SRC>(lanes, sub, [&](int src) {
      float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      int q[3] = {lb[0], lb[1], lb[2]}, tl[3] = {0, 0, 0};
      bool ok = true;
#pragma unroll
      for (int a = 0; a < D; a++) {
        const int d = ((src >> a) & 1) ? (l[a] == kBlock - 1 ? -1 : 1) : 0;
        tl[a] = l[a] + 1 + kBlock * d;
        q[a] = lb[a] - d;
        ok &= (q[a] >= 0) & (q[a] < scenes[0].bn[a]);
      }
      if (ok) {
        const int nb = encode_local<D>(q, P);
        const int c0 = pchunk_begin[nb], cn = pchunk_cnt[nb];
        const int tn = D == 3 ? tl[0] + E8 * (tl[1] + E8 * tl[2]) : tl[0] + E8 * tl[1];
        for (int k = c0; k < c0 + cn; k++) {
          const float4 tv = tiles[(int64_t)k * TN8 + tn];
          acc.x += tv.x;
          acc.y += tv.y;
          if (D == 3) acc.z += tv.z;
          acc.w += tv.w;
        }
      }
      return acc;
    });
    if (sub != 0) continue;
    float m = sum.w, p[3] = {sum.x, sum.y, sum.z};
    if (node_has_mass(m)) {  // mpm_solver.py:591-593
      const float inv = 1.0f / m;
#pragma unroll
      for (int a = 0; a < D; a++) p[a] = inv * p[a];
#pragma unroll
      for (int a = 0; a < D; a++) p[a] += P.dt * P.gravity[a];
    }
    if (P.v_allowed > 0.0f) {  // mpm_solver.py:596-598
#pragma unroll
      for (int a = 0; a < D; a++) p[a] = fminf(fmaxf(p[a], -P.v_allowed), P.v_allowed);
    }
    int I[3], lbl[3];
    to_local<D>(lb, P, lbl);
#pragma unroll
    for (int a = 0; a < D; a++) I[a] = (lbl[a] + P.blo[a]) * kBlock + l[a];
    apply_post<D>(P, I, p);
    grid[(int64_t)bid * NPB + node] = make_float4(p[0], p[1], D == 3 ? p[2] : 0.0f, m);
  }
}

}  // namespace mpm
