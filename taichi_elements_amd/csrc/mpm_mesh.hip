// Mesh seeding on the GPU: the voxelizer of engine/voxelizer.py and the
// particle seeding of MPMSolver.seed_from_voxels (mpm_solver.py:1017-1047).
//
// Voxelization.  One thread per triangle applies the reference's
// column-fill rule (engine/voxelizer.py:45-121, float64, no contraction):
// every column (p, q) whose centre lies in the triangle's xy projection gets
// +1 (normal z > 0) or -1 from z = padding up to the triangle's height.  The
// kernel records each fill as an event at that height (an integer atomic
// add, so the result does not depend on the order), and one thread per
// column turns the events into voxel counts (voxel z counts the fills whose
// height exceeds z).  Heights are capped at res[2] (the reference caps at
// res[1] - padding only, which equals or exceeds it on its cubic grids).
//
// Seeding.  A voxel with a positive count seeds particle l (0 <= l <=
// sample_density) when r + l < sample_density / super_sample^3, at
// ((r0, r1, r2) + (i, j, k)) * voxel_dx + translation -- the reference's
// rule, with r drawn from a counter-based hash of (seed, voxel, l, component)
// instead of ti.random.  One pass counts the particles per column, the host
// scans the counts, and a second pass writes them in column order.
#include <hip/hip_runtime.h>

#include "mpm_kernels.h"

namespace mpm {

namespace {

__device__ __forceinline__ double cross2d(double ax, double ay, double bx, double by) {
#pragma clang fp contract(off)
  return ax * by - ay * bx;
}

__device__ __forceinline__ bool inside_ccw(double px, double py, const double a[2],
                                           const double b[2], const double c[2]) {
#pragma clang fp contract(off)
  return cross2d(a[0] - px, a[1] - py, b[0] - px, b[1] - py) >= 0.0 &&
         cross2d(b[0] - px, b[1] - py, c[0] - px, c[1] - py) >= 0.0 &&
         cross2d(c[0] - px, c[1] - py, a[0] - px, a[1] - py) >= 0.0;
}

// float64 -> int truncation toward zero, saturated so that far-away
// coordinates stay defined (they clamp to the grid below anyway)
__device__ __forceinline__ int to_int(double x) {
  return (int)fmin(fmax(x, -1e9), 1e9);
}

// events[(p * vy + q) * (vz + 1) + h] += inc for every filled column
__global__ __launch_bounds__(kThreads) void k_vox_events(const double* __restrict__ tri,
                                                         int64_t ntri, VoxGrid g,
                                                         int* __restrict__ events) {
#pragma clang fp contract(off)
  // engine/voxelizer.py:56-61: jitter against samples on triangle edges
  const double jit[3] = {-0.057616723909439505 * 1e-8, -0.25608986292614977 * 1e-8,
                         0.06716309129743714 * 1e-8};
  for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < ntri;
       t += (int64_t)gridDim.x * kThreads) {
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; k++) {
      a[k] = tri[t * 9 + k] + jit[k];
      b[k] = tri[t * 9 + 3 + k] + jit[k];
      c[k] = tri[t * 9 + 6 + k] + jit[k];
    }
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(fmin(a[k], b[k]), c[k]);
      hi[k] = fmax(fmax(a[k], b[k]), c[k]);
    }
    const int p_min = max(g.padding, to_int(floor(lo[0] * g.inv_dx)));
    const int p_max = min(g.res[0] - g.padding, to_int(floor(hi[0] * g.inv_dx)) + 1);
    const int q_min = max(g.padding, to_int(floor(lo[1] * g.inv_dx)));
    const int q_max = min(g.res[1] - g.padding, to_int(floor(hi[1] * g.inv_dx)) + 1);
    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                   u[0] * w[1] - u[1] * w[0]};
    const double inv_len = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; k++) n[k] = n[k] * inv_len;
    if (fabs(n[2]) < 1e-10) continue;
    const int inc = n[2] > 0.0 ? 1 : -1;
    const int top = min(g.res[1] - g.padding, g.res[2]);  // reference caps at res[1]
    for (int p = p_min; p < p_max; p++)
      for (int q = q_min; q < q_max; q++) {
        const double px = (p + 0.5) * g.dx, py = (q + 0.5) * g.dx;
        if (!(inside_ccw(px, py, a, b, c) || inside_ccw(px, py, a, c, b))) continue;
        const double d[3] = {px - a[0], py - a[1], 0.0 - a[2]};
        const double dot = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        int h = to_int(-dot / n[2] * g.inv_dx + 0.5);  // truncation toward zero
        h = min(h, top);
        if (h > g.padding)
          atomicAdd(&events[((int64_t)p * g.res[1] + q) * (g.res[2] + 1) + h], inc);
      }
  }
}

// lowbias32 (C. Wellons' integer hash): 32-bit avalanche mixing
__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Uniform float in [0, 1) on a 2^-24 grid from (seed, voxel index, l, comp):
// the four words go through the hash one after another.
__device__ __forceinline__ float uniform01(unsigned seed, uint64_t voxel, unsigned l,
                                           unsigned comp) {
  unsigned h = mix32(l * 4u + comp + 0x9e3779b9u);
  h = mix32(h ^ (unsigned)voxel);
  h = mix32(h ^ (unsigned)(voxel >> 32));
  h = mix32(h ^ seed);
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// Per column: voxel counts from the events (written to voxels when given),
// the particles the column seeds (counts), and, with out_x set, their
// positions from the column's offset.
__global__ __launch_bounds__(kThreads) void k_vox_columns(VoxGrid g, const int* __restrict__ events,
                                                          int* __restrict__ voxels,
                                                          SeedParams sp,
                                                          int* __restrict__ counts,
                                                          const int64_t* __restrict__ offsets,
                                                          float* __restrict__ out_x,
                                                          int64_t out_cap) {
#pragma clang fp contract(off)
  const int64_t ncol = (int64_t)g.res[0] * g.res[1];
  const float s = (float)sp.sample_density / (float)sp.ss3;  // mpm_solver.py:1030
  for (int64_t col = (int64_t)blockIdx.x * kThreads + threadIdx.x; col < ncol;
       col += (int64_t)gridDim.x * kThreads) {
    const int p = (int)(col / g.res[1]), q = (int)(col % g.res[1]);
    const int* ev = events + col * (g.res[2] + 1);
    // voxel z counts the fills whose height h exceeds z: total - sum_{h <= z}
    int total = 0;
    for (int h = 0; h <= g.res[2]; h++) total += ev[h];
    int below = 0, seeded = 0;
    int64_t at = out_x ? offsets[col] : 0;
    // mpm_solver.py:1026-1029 tests the first index only, against the
    // sparse grid's half extent
    const bool seeds = sp.sample_density >= 0 && p < sp.p_limit;
    for (int z = 0; z < g.res[2]; z++) {
      below += ev[z];
      const int v = z >= g.padding ? total - below : 0;
      if (voxels) voxels[col * g.res[2] + z] = v;
      if (v <= 0 || !seeds) continue;
      const uint64_t vid = (uint64_t)col * g.res[2] + z;
      for (int l = 0; l <= sp.sample_density; l++) {
        if ((float)l >= s) break;  // r >= 0: r + l < s fails for this l and every later one
        if (uniform01(sp.seed, vid, (unsigned)l, 3u) + (float)l < s) {  // mpm_solver.py:1032
          if (out_x && at < out_cap) {
            const int ijk[3] = {p, q, z};
            for (int a = 0; a < 3; a++) {
              const float r = uniform01(sp.seed, vid, (unsigned)l, (unsigned)a);
              out_x[at * 3 + a] = (r + (float)ijk[a]) * sp.pdx + sp.shift[a];  // mpm:1033-1038
            }
            at++;
          }
          seeded++;
        }
      }
    }
    if (counts) counts[col] = seeded;
  }
}

__global__ __launch_bounds__(kThreads) void k_fill_seed_attrs(int64_t n, float vx, float vy,
                                                              float vz, int mat,
                                                              float* __restrict__ v,
                                                              int* __restrict__ m) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    v[i * 3 + 0] = vx;
    v[i * 3 + 1] = vy;
    v[i * 3 + 2] = vz;
    m[i] = mat;
  }
}

int groups_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 16384));
}

}  // namespace

void launch_vox_events(const double* tri, int64_t ntri, const VoxGrid& g, int* events,
                       hipStream_t st) {
  if (ntri > 0) k_vox_events<<<groups_for(ntri), kThreads, 0, st>>>(tri, ntri, g, events);
}

void launch_vox_columns(const VoxGrid& g, const int* events, int* voxels, const SeedParams& sp,
                        int* counts, const int64_t* offsets, float* out_x, int64_t out_cap,
                        hipStream_t st) {
  const int64_t ncol = (int64_t)g.res[0] * g.res[1];
  k_vox_columns<<<groups_for(ncol), kThreads, 0, st>>>(g, events, voxels, sp, counts, offsets,
                                                        out_x, out_cap);
}

void launch_seed_attrs(int64_t n, const float v[3], int mat, float* vout, int* mout,
                       hipStream_t st) {
  if (n > 0)
    k_fill_seed_attrs<<<groups_for(n), kThreads, 0, st>>>(n, v[0], v[1], v[2], mat, vout, mout);
}

}  // namespace mpm
