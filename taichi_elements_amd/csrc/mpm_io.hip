// Particle output on the GPU: the quantised particle file of
// ParticleIO.write_particles (engine/particle_io.py:12-76).
//
// Each x and v component is mapped linearly from its [min, max] range onto
// 24 (x) or 8 (v) bits and both are packed into one uint32 per component,
// x in the high bits: x_and_v = (qx << 8) + qv.  The reference evaluates
//   q = uint32(((c - lo) * (1 / (hi - lo))) * (2^bits - 1) + 0.499)
// in float32, one rounding per operation; the kernels do the same with FMA
// contraction disabled, so the output is bit-identical.
#include <hip/hip_runtime.h>

#include "mpm_kernels.h"

namespace mpm {

namespace {

// Per-workgroup minimum and maximum of the 2D components x_0..x_{D-1},
// v_0..v_{D-1} over slots [0, n): partial[(g * 2D + c) * 2 + {0, 1}].
template <int D>
__global__ __launch_bounds__(kThreads) void k_minmax(const float* __restrict__ pf, int64_t cap,
                                                     int n, float* __restrict__ partial) {
  using L = Layout<D>;
  constexpr int NC = 2 * D;
  float lo[NC], hi[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    lo[c] = INFINITY;
    hi[c] = -INFINITY;
  }
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int comp = c < D ? L::X + c : L::V + (c - D);
      const float val = pf[(int64_t)comp * cap + j];
      lo[c] = fminf(lo[c], val);
      hi[c] = fmaxf(hi[c], val);
    }
  }
  __shared__ float red[kThreads / 64][NC][2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < NC; c++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo[c] = fminf(lo[c], __shfl_xor(lo[c], off));
      hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], off));
    }
    if (lane == 0) {
      red[wave][c][0] = lo[c];
      red[wave][c][1] = hi[c];
    }
  }
  __syncthreads();
  if (threadIdx.x < NC * 2) {
    const int c = threadIdx.x >> 1, k = threadIdx.x & 1;
    float r = red[0][c][k];
#pragma unroll
    for (int w = 1; w < kThreads / 64; w++) r = k ? fmaxf(r, red[w][c][k]) : fminf(r, red[w][c][k]);
    partial[((int64_t)blockIdx.x * NC + c) * 2 + k] = r;
  }
}

__device__ __forceinline__ unsigned quantize(float c, float lo, float inv, float levels) {
#pragma clang fp contract(off)
  const float t = (c - lo) * inv;
  const float q = t * levels + 0.499f;
  return (unsigned)q;
}

// x_and_v[orig * D + d] for every slot, in insertion order.
template <int D>
__global__ __launch_bounds__(kThreads) void k_quantize(const float* __restrict__ pf,
                                                       const int* __restrict__ orig, int64_t cap,
                                                       int n, QuantRanges q,
                                                       unsigned* __restrict__ out) {
  using L = Layout<D>;
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
    const int o = orig[j];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const unsigned qx =
          quantize(pf[(int64_t)(L::X + d) * cap + j], q.lo[d], q.inv[d], 16777215.0f);
      const unsigned qv =
          quantize(pf[(int64_t)(L::V + d) * cap + j], q.lo[D + d], q.inv[D + d], 255.0f);
      out[(int64_t)o * D + d] = (qx << 8) + qv;  // uint32 arithmetic, as numpy's
    }
  }
}

}  // namespace

void launch_minmax(int D, const float* pf, int64_t cap, int n, float* partial, int groups,
                   hipStream_t st) {
  if (D == 3)
    k_minmax<3><<<groups, kThreads, 0, st>>>(pf, cap, n, partial);
  else
    k_minmax<2><<<groups, kThreads, 0, st>>>(pf, cap, n, partial);
}

void launch_quantize(int D, const float* pf, const int* orig, int64_t cap, int n,
                     const QuantRanges& q, unsigned* out, hipStream_t st) {
  const int groups = (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads,
                                                                  8192));
  if (D == 3)
    k_quantize<3><<<groups, kThreads, 0, st>>>(pf, orig, cap, n, q, out);
  else
    k_quantize<2><<<groups, kThreads, 0, st>>>(pf, orig, cap, n, q, out);
}

}  // namespace mpm
